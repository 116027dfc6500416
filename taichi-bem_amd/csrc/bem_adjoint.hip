// bem_adjoint.hip -- adjoint contractions of the Galerkin operators with
// respect to the vertex positions and the wave number.
//
// For coefficient vectors lam (test dofs) and u (trial dofs) the kernels add
//
//     grad[v][d] += Re( alpha sum_{s,t} lam_s u_t dA_st / dX_{v,d} )
//     grad_k     += Re( alpha sum_{s,t} lam_s u_t dA_st / dk )
//
// without forming dA: every quadrature point pair of every panel pair is
// evaluated again and its derivative contracted on the spot.  The discrete
// operator differentiated is exactly the one bem_assemble builds (same
// tensor rule for regular pairs, same Sauter-Schwab tables, vertex-slot
// permutations and coincident passes for the near field), so the result is
// the exact derivative of the discrete functional.
//
// Supported: V (P0 x P0) and K (P0 test x P1 trial) -- the operators of the
// Dirichlet problem V t = (M/2 + K) g (BEM_manager.py:990-1011).  Entry
// derivatives restated from the integrands of single_layer_3d.py:155-177
// (G) and double_layer_3d.py:158-180 (dG/dn_y):
//
//   V_ij = sum w_a w_b (2A_i)(2A_j) G(x_a - y_b)
//   K_iv = sum w_a w_b (2A_i) phi_v(y_b) (x_a - y_b).N_j H(|x_a - y_b|)
//
// with N_j = (Y1-Y0) x (Y2-Y0) = 2A_j n_j, 4 pi G = exp(i kap r)/r,
// 4 pi H = (1 - i kap r) exp(i kap r)/r^3, and x_a = sum_m phi_m(a) X_m.
#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

namespace {

struct AdjParams {
    int nf, nv, q2, flags;
    const int32_t* __restrict__ tris;
    const double* __restrict__ verts;
    const double* __restrict__ area;
    const double* __restrict__ normal;
    const double* __restrict__ pu;    // [nf][q2][4] absolute points, weight w (2A)
    const double* __restrict__ pl;    // [q2][4][nf]
    const double* __restrict__ phi;   // [q2][3]
    const NearPair* __restrict__ near;
    const NearPt* __restrict__ rule;
    int32_t rule_off[4], rule_len[4];
    int64_t n_near;
    double kap;    // wave number inside G
    double dkap;   // d kap / d k
    const double* __restrict__ lam;  // complex128 [test dofs]
    const double* __restrict__ u;    // complex128 [trial dofs]
    double ar, ai;                   // alpha / (4 pi) folded in on the host
    double* grad;                    // [nv*3]
    double* gk;                      // [1] or null
    double* kt;                      // K only: complex128 [nv] += (K^T lam) of the regular pairs, or null
    // per-slot partials, each written by exactly one workgroup or wave and
    // summed in a fixed order by the reduction kernels below
    int ti;                          // test panels per regular-pair workgroup
    int nyc;                         // trial-side components: 9, or 15 with the K^T lam terms
    double* fx;                      // [njb][9][nf]   test-panel corners, per trial block
    double* fy;                      // [nib][nyc][nf] trial-panel corners (+ K^T lam), per test chunk
    double* gkb;                     // [nib * njb]    k-derivative, per workgroup
    double* nb;                      // [n_near][NB]   near pairs: test corners, trial corners, k
    double* pc;                      // [nyc][nf]      panel-corner totals
    const int32_t* __restrict__ near_row;
    const int32_t* __restrict__ near_tr;
    const int32_t* __restrict__ vp_row;
    const int32_t* __restrict__ vp_el;
};

constexpr int NB = 20;  // per near pair: 9 test-corner, 9 trial-corner, 1 k-derivative slots (+ pad)

// d(2A)/dX_m / (2A) for the stored slots of panel with vertices P[3][3]:
// 2A = |N|, N = (P1-P0) x (P2-P0): d|N|/dP1 = E2 x n, d|N|/dP2 = n x E1.
__device__ __forceinline__ void dlog_area(const double (&P)[3][3], double (&g)[3][3]) {
    double e1[3], e2[3];
    for (int d = 0; d < 3; ++d) { e1[d] = P[1][d] - P[0][d]; e2[d] = P[2][d] - P[0][d]; }
    double N[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double l2 = N[0] * N[0] + N[1] * N[1] + N[2] * N[2];
    const double il2 = 1.0 / l2;  // d|N|/dP / |N| = (E x N) / |N|^2
    double a[3] = {(e2[1] * N[2] - e2[2] * N[1]) * il2, (e2[2] * N[0] - e2[0] * N[2]) * il2,
                   (e2[0] * N[1] - e2[1] * N[0]) * il2};
    double b[3] = {(N[1] * e1[2] - N[2] * e1[1]) * il2, (N[2] * e1[0] - N[0] * e1[2]) * il2,
                   (N[0] * e1[1] - N[1] * e1[0]) * il2};
    for (int d = 0; d < 3; ++d) { g[1][d] = a[d]; g[2][d] = b[d]; g[0][d] = -a[d] - b[d]; }
}

__device__ __forceinline__ void load_panel(const AdjParams& p, int i, double (&P)[3][3]) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) P[m][d] = p.verts[3 * p.tris[3 * i + m] + d];
}

__device__ __forceinline__ bool shares_vertex_adj(const int32_t* t, int i, int j) {
    int a0 = t[3 * i], a1 = t[3 * i + 1], a2 = t[3 * i + 2];
    int b0 = t[3 * j], b1 = t[3 * j + 1], b2 = t[3 * j + 2];
    return a0 == b0 || a0 == b1 || a0 == b2 || a1 == b0 || a1 == b1 || a1 == b2 || a2 == b0 ||
           a2 == b1 || a2 == b2;
}

// complex helpers on (re, im) pairs
__device__ __forceinline__ void cmac(double& zr, double& zi, double ar, double ai, double br, double bi) {
    zr = __builtin_fma(ar, br, __builtin_fma(-ai, bi, zr));
    zi = __builtin_fma(ar, bi, __builtin_fma(ai, br, zi));
}
__device__ __forceinline__ double cre(double ar, double ai, double br, double bi) { return ar * br - ai * bi; }

// Kernel pieces at separation d (|d| = r), all times 4 pi:
//   V: val = e/r, grd = e (i kap r - 1)/r^3 (so grad_d G = grd d), dk = i e
//   K: h = (1 - i kap r) e/r^3, hp = e (kap^2 r^2 - 3 + 3 i kap r)/r^5,
//      dk(h) = kap e / r
// with e = exp(i kap r).
template <bool HELM>
__device__ __forceinline__ void trig(double kap, double r, double& cs, double& sn) {
    if constexpr (HELM) Math<double>::sincos(kap * r, &sn, &cs);
    else { cs = 1.0; sn = 0.0; }
}
// The same from a block's LDS copy of the sin/cos table.
template <bool HELM>
__device__ __forceinline__ void trig_lds(double kap, double r, const double2* tab, double& cs, double& sn) {
    if constexpr (HELM) sincos_f64_tab(kap * r, tab, &sn, &cs);
    else { cs = 1.0; sn = 0.0; }
}

// ---------------------------------------------------------------------------
// Regular pairs.  Block: 256 trial panels (one per lane) x ti test panels
// (wave-uniform: every lane reads the same test points).  The trial-vertex
// gradient stays in registers over the ti test panels and is stored once per
// (test chunk, trial panel); the test-vertex gradient is reduced over the
// block per test panel and stored once per (trial block, test panel).  No
// atomics: adjoint_panel_reduce_kernel sums the slots in a fixed order, so the
// result is bitwise reproducible.  Launch bounds ask for two waves per SIMD;
// the K Helmholtz variant then spills a few registers, whose loads and stores
// run once per panel pair, outside the point loops.
//
// V is symmetric pair by pair (V_ji sums the same point pairs as V_ij with
// the roles of the two rule copies exchanged, and G(-d) = G(d)), so V runs
// over the pairs i < j only, contracted with lam_i u_j + lam_j u_i; blocks
// with no such pair write zeros to their slots and return.  K runs over all
// ordered pairs.
// ---------------------------------------------------------------------------
__device__ void adjoint_zero_slots(const AdjParams& p, int ib, int jb) {
    const int nf = p.nf, ti = p.ti;
    for (int e = threadIdx.x; e < ti * 9; e += 256) {
        const int i = ib * ti + e / 9;
        if (i < nf) p.fx[((int64_t)jb * 9 + e % 9) * nf + i] = 0.0;
    }
    const int j = jb * 256 + threadIdx.x;
    if (j < nf)
        for (int c = 0; c < p.nyc; ++c) p.fy[((int64_t)ib * p.nyc + c) * nf + j] = 0.0;
    if (threadIdx.x == 0) p.gkb[blockIdx.x] = 0.0;
}

// NQ2 > 0 fixes the rule size so the point loops unroll (their scalar loads
// of the test points are then issued in batches).
template <int OP, bool HELM, int NQ2 = 0>
__global__ __launch_bounds__(256, 2) void adjoint_far_kernel(AdjParams p, int njb) {
    const int ib = blockIdx.x / njb, jb = blockIdx.x % njb;
    const int TI = p.ti;
    if constexpr (OP == BEM_OP_V)
        if (ib * TI >= min(p.nf, (jb + 1) * 256) - 1) {  // every i >= every j
            adjoint_zero_slots(p, ib, jb);
            return;
        }
    __shared__ double2 sct[256];  // sin/cos(j pi/128)
    if constexpr (HELM) {
        sincos_table_fill(sct);
        __syncthreads();
    }
    const int j = jb * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const bool jvalid = j < p.nf;
    const int jj = jvalid ? j : 0;
    const int q2 = NQ2 > 0 ? NQ2 : p.q2, nf = p.nf;
    const double* __restrict__ PL = p.pl;
    const ConstPtr<double> PHI = as_const(p.phi);  // wave-uniform: scalar loads
    const double kap = p.kap;

    // the trial panel's vertices are reloaded per pair in the epilogue
    // rather than held in registers over the point loops
    const double nj[3] = {p.normal[3 * jj], p.normal[3 * jj + 1], p.normal[3 * jj + 2]};
    // trial coefficients: u_j (V) or u at the three vertices (K)
    double ur[3], ui[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int idx = OP == BEM_OP_V ? jj : p.tris[3 * jj + t];
        ur[t] = p.u[2 * idx]; ui[t] = p.u[2 * idx + 1];
    }
    // V: lam_j for the mirrored order (j, i)
    double ljr = 0.0, lji = 0.0;
    if constexpr (OP == BEM_OP_V) { ljr = p.lam[2 * jj]; lji = p.lam[2 * jj + 1]; }
    double gy[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) gy[m][d] = 0.0;
    double gk_lane = 0.0;
    // K: (K^T lam) at the trial panel's three vertices, over the block's test panels
    double ktr[3] = {0.0, 0.0, 0.0}, kti[3] = {0.0, 0.0, 0.0};

    __shared__ double red[4][10];
    const int i_end = min(nf, (ib + 1) * TI);
    for (int i = ib * TI; i < i_end; ++i) {
        double gx[3][3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) gx[m][d] = 0.0;
        double gkp = 0.0;
        if (jvalid && (OP != BEM_OP_V || i < j) && !shares_vertex_adj(p.tris, i, j)) {
            const ConstPtr<double> PU = as_const(p.pu + (int64_t)i * q2 * 4);  // wave-uniform test points
            // accumulators over the pair (complex): test-slot gradient EX,
            // trial-slot gradient EY, value S0, k-derivative SK, and for K the
            // N_j-derivative vector T.  EX[0] is not accumulated: the six
            // point-map gradients sum to zero (d = x - y, sum_m phi_m = 1),
            // so EX[0] = -(EY[0] + EY[1] + EY[2]) - EX[1] - EX[2] after the
            // loops, which keeps the K kernel within two waves per SIMD.
            double EXr[3][3], EXi[3][3], EYr[3][3], EYi[3][3];
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int d = 0; d < 3; ++d) { EXr[m][d] = EXi[m][d] = EYr[m][d] = EYi[m][d] = 0.0; }
            double S0r = 0.0, S0i = 0.0, SKr = 0.0, SKi = 0.0;
            double Tr[3] = {0.0, 0.0, 0.0}, Ti[3] = {0.0, 0.0, 0.0};
            // K: the pair's entries K_iv (times 4 pi) for the trial vertices t;
            // the value sum is then S0 = sum_t u_t KB_t
            double KBr[3] = {0.0, 0.0, 0.0}, KBi[3] = {0.0, 0.0, 0.0};
            // test panel (wave-uniform): corner X0 and edges E1, E2 in stored slot order
            double X0[3], E1[3], E2[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                X0[d] = p.verts[3 * p.tris[3 * i] + d];
                E1[d] = p.verts[3 * p.tris[3 * i + 1] + d] - X0[d];
                E2[d] = p.verts[3 * p.tris[3 * i + 2] + d] - X0[d];
            }
            for (int b = 0; b < q2; ++b) {
                const double yx = PL[((int64_t)b * 4 + 0) * nf + j];
                const double yy = PL[((int64_t)b * 4 + 1) * nf + j];
                const double yz = PL[((int64_t)b * 4 + 2) * nf + j];
                const double wb = PL[((int64_t)b * 4 + 3) * nf + j];
                // P[0] = sum_a grad (all slots), P[m] = sum_a phi_m(a) grad, m = 1, 2
                double Pr[3][3], Pi[3][3];
                double s0r = 0.0, s0i = 0.0, skr = 0.0, ski = 0.0;
                double tr[3] = {0.0, 0.0, 0.0}, ti[3] = {0.0, 0.0, 0.0};
                if constexpr (OP == BEM_OP_V) {
                    // V: grad_d of the integrand is wh d with wh = wa (i kap r - 1) e/r^3;
                    // the phi-weighted sums follow from the moments as for K below
                    double Gr = 0.0, Gi = 0.0, G1r = 0.0, G1i = 0.0, G2r = 0.0, G2i = 0.0;
                    double G11r = 0.0, G11i = 0.0, G12r = 0.0, G12i = 0.0, G22r = 0.0, G22i = 0.0;
#pragma unroll 5
                    for (int a = 0; a < q2; ++a) {
                        const double dv[3] = {PU[4 * a + 0] - yx, PU[4 * a + 1] - yy, PU[4 * a + 2] - yz};
                        const double wa = PU[4 * a + 3];
                        const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
                        const double rinv = rsqrt_f64(r2);
                        const double r = r2 * rinv;
                        double cs, sn;
                        trig_lds<HELM>(kap, r, sct, cs, sn);
                        const double kr = kap * r;
                        const double ri3 = rinv * rinv * rinv;
                        s0r += wa * cs * rinv; s0i += wa * sn * rinv;
                        if constexpr (HELM) { skr -= wa * sn; ski += wa * cs; }
                        const double whr = wa * (-cs - kr * sn) * ri3, whi = wa * (kr * cs - sn) * ri3;
                        const double p1 = PHI[3 * a + 1], p2 = PHI[3 * a + 2];
                        Gr += whr; Gi += whi;
                        const double t1r = p1 * whr, t1i = p1 * whi;
                        G1r += t1r; G1i += t1i;
                        G11r += p1 * t1r; G11i += p1 * t1i;
                        G12r += p2 * t1r; G12i += p2 * t1i;
                        const double t2r = p2 * whr, t2i = p2 * whi;
                        G2r += t2r; G2i += t2i;
                        G22r += p2 * t2r; G22i += p2 * t2i;
                    }
                    const double Dx[3] = {X0[0] - yx, X0[1] - yy, X0[2] - yz};
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        Pr[0][d] = Dx[d] * Gr + E1[d] * G1r + E2[d] * G2r;
                        Pi[0][d] = Dx[d] * Gi + E1[d] * G1i + E2[d] * G2i;
                        Pr[1][d] = Dx[d] * G1r + E1[d] * G11r + E2[d] * G12r;
                        Pi[1][d] = Dx[d] * G1i + E1[d] * G11i + E2[d] * G12i;
                        Pr[2][d] = Dx[d] * G2r + E1[d] * G12r + E2[d] * G22r;
                        Pi[2][d] = Dx[d] * G2i + E1[d] * G12i + E2[d] * G22i;
                    }
                } else {
                    // K: grad_d of the integrand is wh n_j + wp d with
                    // wh = wa h, wp = wa (d.n_j) hp.  Since d = x_a - y_b =
                    // (X0 - y_b) + phi_1(a) E1 + phi_2(a) E2, the sums over a
                    // of grad, phi_1 grad, phi_2 grad and T = sum wh d follow
                    // from the complex moments sum {1, phi_1, phi_2} wh and
                    // sum {1, phi_1, phi_2, phi_1^2, phi_1 phi_2, phi_2^2} wp,
                    // with no 3-vector accumulated per point pair.
                    double Hr = 0.0, Hi = 0.0, H1r = 0.0, H1i = 0.0, H2r = 0.0, H2i = 0.0;
                    double Gr = 0.0, Gi = 0.0, G1r = 0.0, G1i = 0.0, G2r = 0.0, G2i = 0.0;
                    double G11r = 0.0, G11i = 0.0, G12r = 0.0, G12i = 0.0, G22r = 0.0, G22i = 0.0;
#pragma unroll 5
                    for (int a = 0; a < q2; ++a) {
                        const double dv[3] = {PU[4 * a + 0] - yx, PU[4 * a + 1] - yy, PU[4 * a + 2] - yz};
                        const double wa = PU[4 * a + 3];
                        const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
                        const double rinv = rsqrt_f64(r2);
                        const double r = r2 * rinv;
                        double cs, sn;
                        trig_lds<HELM>(kap, r, sct, cs, sn);
                        const double kr = kap * r;
                        const double ri3 = rinv * rinv * rinv;
                        const double dn = dv[0] * nj[0] + dv[1] * nj[1] + dv[2] * nj[2];
                        const double hr = (cs + kr * sn) * ri3, hi = (sn - kr * cs) * ri3;
                        const double ri5 = ri3 * rinv * rinv;
                        const double A = kr * kr - 3.0, B = 3.0 * kr;
                        const double hpr = (cs * A - sn * B) * ri5, hpi = (sn * A + cs * B) * ri5;
                        const double wd = wa * dn;
                        s0r += wd * hr; s0i += wd * hi;
                        if constexpr (HELM) {
                            const double f = wd * kap * rinv;
                            skr += f * cs; ski += f * sn;
                        }
                        const double whr = wa * hr, whi = wa * hi;
                        const double wpr = wd * hpr, wpi = wd * hpi;
                        const double p1 = PHI[3 * a + 1], p2 = PHI[3 * a + 2];
                        Hr += whr; Hi += whi;
                        H1r += p1 * whr; H1i += p1 * whi;
                        H2r += p2 * whr; H2i += p2 * whi;
                        Gr += wpr; Gi += wpi;
                        const double t1r = p1 * wpr, t1i = p1 * wpi;
                        G1r += t1r; G1i += t1i;
                        G11r += p1 * t1r; G11i += p1 * t1i;
                        G12r += p2 * t1r; G12i += p2 * t1i;
                        const double t2r = p2 * wpr, t2i = p2 * wpi;
                        G2r += t2r; G2i += t2i;
                        G22r += p2 * t2r; G22i += p2 * t2i;
                    }
                    const double Dx[3] = {X0[0] - yx, X0[1] - yy, X0[2] - yz};
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        Pr[0][d] = nj[d] * Hr + Dx[d] * Gr + E1[d] * G1r + E2[d] * G2r;
                        Pi[0][d] = nj[d] * Hi + Dx[d] * Gi + E1[d] * G1i + E2[d] * G2i;
                        Pr[1][d] = nj[d] * H1r + Dx[d] * G1r + E1[d] * G11r + E2[d] * G12r;
                        Pi[1][d] = nj[d] * H1i + Dx[d] * G1i + E1[d] * G11i + E2[d] * G12i;
                        Pr[2][d] = nj[d] * H2r + Dx[d] * G2r + E1[d] * G12r + E2[d] * G22r;
                        Pi[2][d] = nj[d] * H2i + Dx[d] * G2i + E1[d] * G12i + E2[d] * G22i;
                        tr[d] = Dx[d] * Hr + E1[d] * H1r + E2[d] * H2r;
                        ti[d] = Dx[d] * Hi + E1[d] * H1i + E2[d] * H2i;
                    }
                }
                // trial-point factor: wb (V) or wb * u interpolated at y_b (K)
                double cr = wb, ci = 0.0;
                if constexpr (OP == BEM_OP_K) {
                    const double f0 = PHI[3 * b], f1 = PHI[3 * b + 1], f2 = PHI[3 * b + 2];
                    cr = wb * (f0 * ur[0] + f1 * ur[1] + f2 * ur[2]);
                    ci = wb * (f0 * ui[0] + f1 * ui[1] + f2 * ui[2]);
                }
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double ph = PHI[3 * b + m];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        if (m > 0) cmac(EXr[m][d], EXi[m][d], cr, ci, Pr[m][d], Pi[m][d]);
                        cmac(EYr[m][d], EYi[m][d], -cr * ph, -ci * ph, Pr[0][d], Pi[0][d]);
                    }
                }
                if constexpr (OP == BEM_OP_K) {
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        const double wt = wb * PHI[3 * b + t];
                        KBr[t] += wt * s0r; KBi[t] += wt * s0i;
                    }
                } else {
                    cmac(S0r, S0i, cr, ci, s0r, s0i);
                }
                cmac(SKr, SKi, cr, ci, skr, ski);
                if constexpr (OP == BEM_OP_K) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) cmac(Tr[d], Ti[d], cr, ci, tr[d], ti[d]);
                }
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                EXr[0][d] = -(EYr[0][d] + EYr[1][d] + EYr[2][d]) - EXr[1][d] - EXr[2][d];
                EXi[0][d] = -(EYi[0][d] + EYi[1][d] + EYi[2][d]) - EXi[1][d] - EXi[2][d];
            }
            // contraction coefficient c = alpha lam_i (K) or
            // alpha (lam_i u_j + lam_j u_i) (V)
            double lr = p.lam[2 * i], li = p.lam[2 * i + 1];
            if constexpr (OP == BEM_OP_K) {
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    cmac(S0r, S0i, ur[t], ui[t], KBr[t], KBi[t]);
                    cmac(ktr[t], kti[t], lr, li, KBr[t], KBi[t]);
                }
            }
            if constexpr (OP == BEM_OP_V) {
                const double uir = p.u[2 * i], uii = p.u[2 * i + 1];
                const double t = lr * ur[0] - li * ui[0] + (ljr * uir - lji * uii);
                li = lr * ui[0] + li * ur[0] + (ljr * uii + lji * uir);
                lr = t;
            }
            const double cr = p.ar * lr - p.ai * li, ci = p.ar * li + p.ai * lr;
            double X[3][3], Y[3][3];
            load_panel(p, i, X);
            load_panel(p, j, Y);
            double dli[3][3], dlj[3][3];
            dlog_area(X, dli);
            if constexpr (OP == BEM_OP_V) dlog_area(Y, dlj);
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    gx[m][d] = cre(cr, ci, S0r * dli[m][d] + EXr[m][d], S0i * dli[m][d] + EXi[m][d]);
                    double yr = EYr[m][d], yi = EYi[m][d];
                    if constexpr (OP == BEM_OP_V) { yr += S0r * dlj[m][d]; yi += S0i * dlj[m][d]; }
                    gy[m][d] += cre(cr, ci, yr, yi);
                }
            if constexpr (OP == BEM_OP_K) {
                // d(d.N_j)/dY: slot 1 gets E2 x T~, slot 2 gets T~ x E1 (T~ = T / 2A_j)
                double e1[3], e2[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) { e1[d] = Y[1][d] - Y[0][d]; e2[d] = Y[2][d] - Y[0][d]; }
                double tr_[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) tr_[d] = cre(cr, ci, Tr[d], Ti[d]) / (2.0 * p.area[j]);
                const double c1[3] = {e2[1] * tr_[2] - e2[2] * tr_[1], e2[2] * tr_[0] - e2[0] * tr_[2],
                                      e2[0] * tr_[1] - e2[1] * tr_[0]};
                const double c2[3] = {tr_[1] * e1[2] - tr_[2] * e1[1], tr_[2] * e1[0] - tr_[0] * e1[2],
                                      tr_[0] * e1[1] - tr_[1] * e1[0]};
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    gy[1][d] += c1[d];
                    gy[2][d] += c2[d];
                    gy[0][d] -= c1[d] + c2[d];
                }
            }
            gkp = cre(cr, ci, SKr, SKi) * p.dkap;
        }
        // test-vertex gradient of panel i: block reduction, one slot per trial block
        gk_lane += gkp;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double s = wave_sum(gx[m][d]);
                if (lane == 0) red[wid][3 * m + d] = s;
            }
        __syncthreads();
        if (threadIdx.x < 9) {
            const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
            p.fx[((int64_t)jb * 9 + threadIdx.x) * nf + i] = s;
        }
        __syncthreads();
    }
    if (jvalid) {
        double* fy = p.fy + (int64_t)ib * p.nyc * nf + j;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int d = 0; d < 3; ++d) fy[(int64_t)(3 * m + d) * nf] = gy[m][d];
        if constexpr (OP == BEM_OP_K) {
            if (p.nyc == 15) {
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    fy[(int64_t)(9 + 2 * t) * nf] = kInv4Pi * ktr[t];
                    fy[(int64_t)(10 + 2 * t) * nf] = kInv4Pi * kti[t];
                }
            }
        }
    }
    {
        const double s = wave_sum(gk_lane);
        if (lane == 0) red[wid][9] = s;
        __syncthreads();
        if (threadIdx.x == 0) p.gkb[blockIdx.x] = red[0][9] + red[1][9] + red[2][9] + red[3][9];
    }
}

// ---------------------------------------------------------------------------
// Near pairs: one wave per pair over the Sauter-Schwab points of its
// relation, in the (possibly rotated) slot order of the pair, with the same
// coincident passes as the assembly (f(y, x) in the second pass, bases
// swapped under BEM_FLAG_SYM_COINCIDE).
// ---------------------------------------------------------------------------
template <int OP, bool HELM>
__global__ __launch_bounds__(256) void adjoint_near_kernel(AdjParams p) {
    const int lane = threadIdx.x & 63;
    const int64_t pid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pid >= p.n_near) return;
    const NearPair np = p.near[pid];
    const int i = np.i, j = np.j, rel = np.rel;
    int px[3], py[3], invy[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        px[m] = (np.perm >> (2 * m)) & 3;
        py[m] = (np.perm >> (8 + 2 * m)) & 3;
    }
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) invy[sl] = py[0] == sl ? 0 : (py[1] == sl ? 1 : 2);
    double X[3][3], Y[3][3];  // rotated slot order
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            X[m][d] = p.verts[3 * p.tris[3 * i + px[m]] + d];
            Y[m][d] = p.verts[3 * p.tris[3 * j + py[m]] + d];
        }
    const double nj[3] = {p.normal[3 * j], p.normal[3 * j + 1], p.normal[3 * j + 2]};
    const double w4a = (p.area[i] * 2.0) * (p.area[j] * 2.0);
    const double kap = p.kap;
    const bool twice = rel == BEM_COINCIDE;
    const bool sym = (p.flags & BEM_FLAG_SYM_COINCIDE) != 0;
    const NearPt* __restrict__ rule = p.rule + p.rule_off[rel];
    const int len = p.rule_len[rel];
    // trial coefficients at stored slots
    double ur[3], ui[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int idx = OP == BEM_OP_V ? j : p.tris[3 * j + t];
        ur[t] = p.u[2 * idx]; ui[t] = p.u[2 * idx + 1];
    }

    double EXr[3][3], EXi[3][3], EYr[3][3], EYi[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) { EXr[m][d] = EXi[m][d] = EYr[m][d] = EYi[m][d] = 0.0; }
    double S0r = 0.0, S0i = 0.0, SKr = 0.0, SKi = 0.0;
    double Tr[3] = {0.0, 0.0, 0.0}, Ti[3] = {0.0, 0.0, 0.0};

    for (int e = lane; e < len; e += 64) {
        const NearPt q = rule[e];
        const double al[3] = {1.0 - q.r1x, q.r1x - q.r2x, q.r2x};
        const double be[3] = {1.0 - q.r1y, q.r1y - q.r2y, q.r2y};
        double x[3], y[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            x[d] = al[0] * X[0][d] + al[1] * X[1][d] + al[2] * X[2][d];
            y[d] = be[0] * Y[0][d] + be[1] * Y[1][d] + be[2] * Y[2][d];
        }
        const double w = q.w * w4a;
        for (int pass = 0; pass <= (int)twice; ++pass) {
            const double sg = pass ? -1.0 : 1.0;
            const double dv[3] = {sg * (x[0] - y[0]), sg * (x[1] - y[1]), sg * (x[2] - y[2])};
            const double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            const double rinv = rsqrt_f64(r2);
            const double r = r2 * rinv;
            double cs, sn;
            trig<HELM>(kap, r, cs, sn);
            const double kr = kap * r;
            const double ri3 = rinv * rinv * rinv;
            // point factor: w (V) or w * u at the trial basis point (K)
            double cr = w, ci = 0.0;
            if constexpr (OP == BEM_OP_K) {
                const bool swap_b = pass && sym;
                const double by1 = swap_b ? q.r1x : q.r1y, by2 = swap_b ? q.r2x : q.r2y;
                double sr = 0.0, si = 0.0;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const double ph = shape(by1, by2, invy[t]);
                    sr += ph * ur[t]; si += ph * ui[t];
                }
                cr = w * sr; ci = w * si;
            }
            double gdr[3], gdi[3];
            if constexpr (OP == BEM_OP_V) {
                cmac(S0r, S0i, cr, ci, cs * rinv, sn * rinv);
                if constexpr (HELM) cmac(SKr, SKi, cr, ci, -sn, cs);
                double hr = (-cs - kr * sn) * ri3, hi = (kr * cs - sn) * ri3;
                const double t = cr * hr - ci * hi;
                hi = cr * hi + ci * hr; hr = t;
#pragma unroll
                for (int d = 0; d < 3; ++d) { gdr[d] = hr * dv[d]; gdi[d] = hi * dv[d]; }
            } else {
                const double dn = dv[0] * nj[0] + dv[1] * nj[1] + dv[2] * nj[2];
                double hr = (cs + kr * sn) * ri3, hi = (sn - kr * cs) * ri3;
                const double ri5 = ri3 * rinv * rinv;
                const double A = kr * kr - 3.0, B = 3.0 * kr;
                double hpr = (cs * A - sn * B) * ri5, hpi = (sn * A + cs * B) * ri5;
                cmac(S0r, S0i, cr, ci, dn * hr, dn * hi);
                if constexpr (HELM) cmac(SKr, SKi, cr, ci, dn * kap * rinv * cs, dn * kap * rinv * sn);
                double t = cr * hr - ci * hi;
                hi = cr * hi + ci * hr; hr = t;
                t = cr * hpr - ci * hpi;
                hpi = (cr * hpi + ci * hpr) * dn; hpr = t * dn;
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    gdr[d] = hr * nj[d] + hpr * dv[d];
                    gdi[d] = hi * nj[d] + hpi * dv[d];
                    Tr[d] += hr * dv[d]; Ti[d] += hi * dv[d];
                }
            }
            // d = sg (x - y): dd/dX_m = sg al_m, dd/dY_m = -sg be_m
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double fa = sg * al[m], fb = -sg * be[m];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    EXr[m][d] += fa * gdr[d]; EXi[m][d] += fa * gdi[d];
                    EYr[m][d] += fb * gdr[d]; EYi[m][d] += fb * gdi[d];
                }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            EXr[m][d] = wave_sum(EXr[m][d]); EXi[m][d] = wave_sum(EXi[m][d]);
            EYr[m][d] = wave_sum(EYr[m][d]); EYi[m][d] = wave_sum(EYi[m][d]);
        }
    S0r = wave_sum(S0r); S0i = wave_sum(S0i);
    SKr = wave_sum(SKr); SKi = wave_sum(SKi);
    if constexpr (OP == BEM_OP_K) {
#pragma unroll
        for (int d = 0; d < 3; ++d) { Tr[d] = wave_sum(Tr[d]); Ti[d] = wave_sum(Ti[d]); }
    }
    if (lane != 0) return;
    double lr = p.lam[2 * i], li = p.lam[2 * i + 1];
    if constexpr (OP == BEM_OP_V) {
        const double t = lr * ur[0] - li * ui[0];
        li = lr * ui[0] + li * ur[0];
        lr = t;
    }
    const double cr = p.ar * lr - p.ai * li, ci = p.ar * li + p.ai * lr;
    // area (and, for K, N_j) derivatives in stored slot order
    double Xs[3][3], Ys[3][3], gi[3][3], gj[3][3];
    load_panel(p, i, Xs);
    load_panel(p, j, Ys);
    dlog_area(Xs, gi);
    dlog_area(Ys, gj);
    const double s0 = cre(cr, ci, S0r, S0i);
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            gi[m][d] *= s0;
            gj[m][d] = OP == BEM_OP_V ? gj[m][d] * s0 : 0.0;
        }
    if constexpr (OP == BEM_OP_K) {
        const double a2j = 2.0 * p.area[j];
        double e1[3], e2[3], t[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = Ys[1][d] - Ys[0][d]; e2[d] = Ys[2][d] - Ys[0][d];
            t[d] = cre(cr, ci, Tr[d], Ti[d]) / a2j;
        }
        const double c1[3] = {e2[1] * t[2] - e2[2] * t[1], e2[2] * t[0] - e2[0] * t[2], e2[0] * t[1] - e2[1] * t[0]};
        const double c2[3] = {t[1] * e1[2] - t[2] * e1[1], t[2] * e1[0] - t[0] * e1[2], t[0] * e1[1] - t[1] * e1[0]};
        for (int d = 0; d < 3; ++d) {
            gj[1][d] += c1[d];
            gj[2][d] += c2[d];
            gj[0][d] -= c1[d] + c2[d];
        }
    }
    // rotated-slot point derivatives, moved to the stored slots px[m], py[m]
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int sl = 0; sl < 3; ++sl)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                if (px[m] == sl) gi[sl][d] += cre(cr, ci, EXr[m][d], EXi[m][d]);
                if (py[m] == sl) gj[sl][d] += cre(cr, ci, EYr[m][d], EYi[m][d]);
            }
    double* o = p.nb + pid * NB;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            o[3 * m + d] = gi[m][d];
            o[9 + 3 * m + d] = gj[m][d];
        }
    o[18] = cre(cr, ci, SKr, SKi) * p.dkap;
}

// Panel-corner totals in a fixed order: thread per panel q and component
// c = blockIdx.y (3 slot + dim, then for K^T lam the (re, im) pairs of the
// three slots): the trial-block slots of q as a test panel, the test-chunk
// slots of q as a trial panel, its near pairs as test panel (row q) and as
// trial panel (the mirrors of row q).
__global__ void adjoint_panel_reduce_kernel(AdjParams p, int njb, int nib) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (q >= p.nf) return;
    const int nf = p.nf;
    double s = 0.0;
    if (c < 9)
        for (int jb = 0; jb < njb; ++jb) s += p.fx[((int64_t)jb * 9 + c) * nf + q];
    for (int ib = 0; ib < nib; ++ib) s += p.fy[((int64_t)ib * p.nyc + c) * nf + q];
    if (c < 9 && p.n_near > 0) {
        for (int e = p.near_row[q]; e < p.near_row[q + 1]; ++e) s += p.nb[(int64_t)e * NB + c];
        for (int e = p.near_row[q]; e < p.near_row[q + 1]; ++e) s += p.nb[(int64_t)p.near_tr[e] * NB + 9 + c];
    }
    p.pc[(int64_t)c * nf + q] = s;
}

// Vertex sums over the vertex's panel corners in panel order.
__global__ void adjoint_vertex_kernel(AdjParams p) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= p.nv) return;
    const int nf = p.nf;
    double g[3] = {0.0, 0.0, 0.0}, kr = 0.0, ki = 0.0;
    for (int c = p.vp_row[v]; c < p.vp_row[v + 1]; ++c) {
        const int q = p.vp_el[c] / 3, sl = p.vp_el[c] % 3;
        for (int d = 0; d < 3; ++d) g[d] += p.pc[(int64_t)(3 * sl + d) * nf + q];
        if (p.nyc == 15) {
            kr += p.pc[(int64_t)(9 + 2 * sl) * nf + q];
            ki += p.pc[(int64_t)(10 + 2 * sl) * nf + q];
        }
    }
    for (int d = 0; d < 3; ++d) p.grad[3 * v + d] += g[d];
    if (p.nyc == 15) { p.kt[2 * v] += kr; p.kt[2 * v + 1] += ki; }
}

// k-derivative: one workgroup, a fixed assignment of slots to threads and a
// fixed tree.
__global__ __launch_bounds__(256) void adjoint_gk_kernel(AdjParams p, int64_t nblocks) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += 256) s += p.gkb[b];
    for (int64_t e = threadIdx.x; e < p.n_near; e += 256) s += p.nb[e * NB + 18];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *p.gk += red[0];
}

// grad += sum_i coef_i dA_i/dX: one thread per vertex over its panel corners
// in panel order.
__global__ void area_grad_kernel(int nv, const int32_t* __restrict__ vp_row, const int32_t* __restrict__ vp_el,
                                 const int32_t* __restrict__ tris, const double* __restrict__ verts,
                                 const double* __restrict__ area, const double* __restrict__ coef, double* grad) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int c = vp_row[v]; c < vp_row[v + 1]; ++c) {
        const int i = vp_el[c] / 3, sl = vp_el[c] % 3;
        double P[3][3], g[3][3];
        for (int m = 0; m < 3; ++m)
            for (int d = 0; d < 3; ++d) P[m][d] = verts[3 * tris[3 * i + m] + d];
        dlog_area(P, g);
        const double s = coef[i] * area[i];  // dA = A dlog(2A)
        for (int d = 0; d < 3; ++d) acc[d] += s * g[sl][d];
    }
    for (int d = 0; d < 3; ++d) grad[3 * v + d] += acc[d];
}

template <int OP>
void launch_adjoint(const AdjParams& p, bool helm, int njb, int nib, hipStream_t s) {
    const dim3 gf((unsigned)((int64_t)nib * njb));
    if (p.q2 == 25) {
        if (helm) hipLaunchKernelGGL((adjoint_far_kernel<OP, true, 25>), gf, dim3(256), 0, s, p, njb);
        else hipLaunchKernelGGL((adjoint_far_kernel<OP, false, 25>), gf, dim3(256), 0, s, p, njb);
    } else {
        if (helm) hipLaunchKernelGGL((adjoint_far_kernel<OP, true>), gf, dim3(256), 0, s, p, njb);
        else hipLaunchKernelGGL((adjoint_far_kernel<OP, false>), gf, dim3(256), 0, s, p, njb);
    }
    if (p.n_near > 0) {
        const dim3 gn((unsigned)((p.n_near + 3) / 4));
        if (helm) hipLaunchKernelGGL((adjoint_near_kernel<OP, true>), gn, dim3(256), 0, s, p);
        else hipLaunchKernelGGL((adjoint_near_kernel<OP, false>), gn, dim3(256), 0, s, p);
    }
    hipLaunchKernelGGL(adjoint_panel_reduce_kernel, dim3((p.nf + 255) / 256, p.nyc), dim3(256), 0, s, p, njb, nib);
    hipLaunchKernelGGL(adjoint_vertex_kernel, dim3((p.nv + 255) / 256), dim3(256), 0, s, p);
    if (p.gk) hipLaunchKernelGGL(adjoint_gk_kernel, dim3(1), dim3(256), 0, s, p, (int64_t)nib * njb);
}

// Trial-side slots cap: test chunks grow (ti doubles from 8) until the
// per-chunk partials fit in it.
constexpr size_t ADJ_FY_CAP = size_t(1) << 30;

}  // namespace

int shape_gradient(Context* ctx, const bem_op_desc* d, const void* lam, const void* u, double ar, double ai,
                   double* grad, double* grad_k, void* kt, hipStream_t s) {
    Mesh& m = ctx->mesh;
    if (!m.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
    if (!d || !lam || !u || !grad) return set_error(BEM_ERR_INVALID, "null argument");
    if (d->kernel < BEM_LAPLACE || d->kernel > BEM_HELMHOLTZ_TRANSMISSION)
        return set_error(BEM_ERR_INVALID, "Kernel Type only support Laplace and Helmholtz for now");
    const bool v_op = d->op == BEM_OP_V && d->deg_test == 0 && d->deg_trial == 0;
    const bool k_op = d->op == BEM_OP_K && d->deg_test == 0 && d->deg_trial == 1;
    if (!v_op && !k_op)
        return set_error(BEM_ERR_UNSUPPORTED, "shape gradient: V (P0 x P0) and K (P0 x P1) only");
    if (kt && !k_op) return set_error(BEM_ERR_INVALID, "K^T lam output: the operator must be K (P0 x P1)");
    if (d->dtype != BEM_C128) return set_error(BEM_ERR_INVALID, "shape gradient: coefficients must be complex128");
    if (d->precision != BEM_PREC_FP64) return set_error(BEM_ERR_UNSUPPORTED, "shape gradient runs in FP64 only");
    if (d->flags & (BEM_FLAG_DK | BEM_FLAG_TRANSPOSE))
        return set_error(BEM_ERR_INVALID, "BEM_FLAG_DK and BEM_FLAG_TRANSPOSE are bem_apply flags");
    int rc = mesh_build_near(ctx, d->flags, s);
    if (rc) return rc;
    AdjParams p;
    p.nf = m.nf; p.nv = m.nv; p.q2 = m.q2; p.flags = d->flags;
    p.tris = m.tris; p.verts = m.verts; p.area = m.area; p.normal = m.normal;
    p.pu = m.pu; p.pl = m.pl; p.phi = m.phi; p.near = m.near; p.rule = m.rule;
    for (int r = 0; r < 4; ++r) { p.rule_off[r] = m.rule_off[r]; p.rule_len[r] = m.rule_len[r]; }
    p.n_near = m.n_near;
    const bool helm = d->kernel != BEM_LAPLACE;
    p.kap = d->kernel == BEM_HELMHOLTZ_TRANSMISSION ? d->k * d->sqrt_n : (helm ? d->k : 0.0);
    p.dkap = d->kernel == BEM_HELMHOLTZ_TRANSMISSION ? d->sqrt_n : 1.0;
    p.lam = (const double*)lam; p.u = (const double*)u;
    p.ar = ar * kInv4Pi; p.ai = ai * kInv4Pi;
    p.grad = grad; p.gk = helm ? grad_k : nullptr;
    p.kt = (double*)kt;
    const int njb = (m.nf + 255) / 256;
    p.nyc = kt ? 15 : 9;
    p.ti = 8;
    while (p.ti < m.nf && (size_t)((m.nf + p.ti - 1) / p.ti) * p.nyc * m.nf * sizeof(double) > ADJ_FY_CAP) p.ti *= 2;
    const int nib = (m.nf + p.ti - 1) / p.ti;
    const size_t nfx = (size_t)njb * 9 * m.nf, nfy = (size_t)nib * p.nyc * m.nf, ngk = (size_t)nib * njb;
    const size_t nnb = (size_t)m.n_near * NB, npc = (size_t)p.nyc * m.nf;
    const size_t bytes = (nfx + nfy + ngk + nnb + npc) * sizeof(double);
    if (m.adj_bytes < bytes) {
        if (m.adj) hipFree(m.adj);
        m.adj = nullptr;
        m.adj_bytes = 0;
        rc = check_hip(hipMalloc((void**)&m.adj, bytes), "adjoint partials");
        if (rc) return rc;
        m.adj_bytes = bytes;
    }
    p.fx = m.adj; p.fy = p.fx + nfx; p.gkb = p.fy + nfy; p.nb = p.gkb + ngk; p.pc = p.nb + nnb;
    p.near_row = m.near_row; p.near_tr = m.near_tr; p.vp_row = m.vp_row; p.vp_el = m.vp_el;
    if (v_op) launch_adjoint<BEM_OP_V>(p, helm, njb, nib, s);
    else launch_adjoint<BEM_OP_K>(p, helm, njb, nib, s);
    rc = check_hip(hipGetLastError(), "shape gradient launch");
    if (rc || !kt) return rc;
    // the near pairs of K^T lam: the transposed near-field product
    bem_op_desc dn = *d;
    dn.flags |= BEM_FLAG_TRANSPOSE;
    dn.parts = BEM_PART_NEAR;
    return apply_op(ctx, &dn, lam, 0, kt, 1.0, 0.0, s);
}

int area_gradient(Context* ctx, const double* coef, double* grad, hipStream_t s) {
    Mesh& m = ctx->mesh;
    if (!m.nf) return set_error(BEM_ERR_INVALID, "no mesh: call bem_set_mesh first");
    if (!coef || !grad) return set_error(BEM_ERR_INVALID, "null argument");
    hipLaunchKernelGGL(area_grad_kernel, dim3((m.nv + 255) / 256), dim3(256), 0, s, m.nv,
                       (const int32_t*)m.vp_row, (const int32_t*)m.vp_el, (const int32_t*)m.tris,
                       (const double*)m.verts, (const double*)m.area, coef, grad);
    return check_hip(hipGetLastError(), "area gradient launch");
}

}  // namespace bem
