// bem_device.h -- device-side math shared by the BEM kernels (gfx950).
//
// Kernels (reference src/managers/BEM_manager.py:452-525, 3D branch):
//   G(x,y)      = exp(i kap r) / (4 pi r)
//   dG/dn_y     = (x-y).n_y (1 - i kap r) exp(i kap r) / (4 pi r^3)
// with kap = k (Helmholtz) or k*sqrt(n) (transmission); Laplace drops the
// exponential.  The 1/(4 pi) factor is applied once per entry, not per point.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bem {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;  // 1/(4 pi)

// ---------------------------------------------------------------------------
// sincos for kap*r >= 0 of moderate size (|x| < 1e5).  Cody-Waite reduction
// by the double-double split of pi/2 (fma keeps n*hi exact), then the Taylor
// series of sin (through t^15) and cos (through t^16) on |t| <= pi/4, whose
// truncation error is below 7e-17 there.  Quadrant from n mod 4.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos_f64(double x, double* s_out, double* c_out) {
    const double two_over_pi = 0.63661977236758134308;
    const double pio2_hi = 1.5707963267948966192;     // double nearest pi/2
    const double pio2_lo = 6.1232339957367658e-17;    // pi/2 - pio2_hi
    double n = __builtin_rint(x * two_over_pi);
    double t = __builtin_fma(-n, pio2_hi, x);
    t = __builtin_fma(-n, pio2_lo, t);
    const double t2 = t * t;
    // sin: t + t^3 * P(t^2), coefficients (-1)^k / (2k+1)!
    double ps = -7.6471637318198164759e-13;           // -1/15!
    ps = __builtin_fma(ps, t2, 1.6059043836821614599e-10);   // 1/13!
    ps = __builtin_fma(ps, t2, -2.5052108385441718775e-8);   // -1/11!
    ps = __builtin_fma(ps, t2, 2.7557319223985890653e-6);    // 1/9!
    ps = __builtin_fma(ps, t2, -1.9841269841269841270e-4);   // -1/7!
    ps = __builtin_fma(ps, t2, 8.3333333333333333333e-3);    // 1/5!
    ps = __builtin_fma(ps, t2, -1.6666666666666666667e-1);   // -1/3!
    const double s = __builtin_fma(t * t2, ps, t);
    // cos: 1 + t^2 * Q(t^2), coefficients (-1)^k / (2k)!
    double pc = 4.7794773323873852974e-14;             // 1/16!
    pc = __builtin_fma(pc, t2, -1.1470745597729724714e-11);  // -1/14!
    pc = __builtin_fma(pc, t2, 2.0876756987868098979e-9);    // 1/12!
    pc = __builtin_fma(pc, t2, -2.7557319223985890653e-7);   // -1/10!
    pc = __builtin_fma(pc, t2, 2.4801587301587301587e-5);    // 1/8!
    pc = __builtin_fma(pc, t2, -1.3888888888888888889e-3);   // -1/6!
    pc = __builtin_fma(pc, t2, 4.1666666666666666667e-2);    // 1/4!
    pc = __builtin_fma(pc, t2, -0.5);                        // -1/2!
    const double c = __builtin_fma(t2, pc, 1.0);
    const int q = static_cast<int>(n) & 3;
    const double ss = (q & 1) ? c : s;
    const double cc = (q & 1) ? s : c;
    *s_out = (q & 2) ? -ss : ss;
    *c_out = ((q + 1) & 2) ? -cc : cc;
}

// sincos for x >= 0 from a 64-entry table of sin/cos(j pi/32) (kept in LDS by
// the caller, tab[j] = {sin, cos}): x = n pi/32 + t with |t| <= pi/64 (the
// Cody-Waite split of pi/32 is the pi/2 split above scaled by 1/16, exact),
// Taylor series of sin t through t^7 and cos t through t^8 (truncation below
// 5e-18 on |t| <= pi/64), then the angle-sum formulas with the table entry
// of n mod 64.  About two thirds of the instructions of sincos_f64.
__device__ __forceinline__ void sincos_f64_tab(double x, const double2* tab, double* s_out, double* c_out) {
    const double inv_pio32 = 10.185916357881301489;      // 32 / pi
    const double pio32_hi = 0.098174770424681038702;     // pio2_hi / 16
    const double pio32_lo = 3.8270212473354786e-18;      // pio2_lo / 16
    const double n = __builtin_rint(x * inv_pio32);
    double t = __builtin_fma(-n, pio32_hi, x);
    t = __builtin_fma(-n, pio32_lo, t);
    const double t2 = t * t;
    double ps = -1.9841269841269841270e-4;                    // -1/7!
    ps = __builtin_fma(ps, t2, 8.3333333333333333333e-3);     // 1/5!
    ps = __builtin_fma(ps, t2, -1.6666666666666666667e-1);    // -1/3!
    const double st = __builtin_fma(t * t2, ps, t);
    double pc = 2.4801587301587301587e-5;                     // 1/8!
    pc = __builtin_fma(pc, t2, -1.3888888888888888889e-3);    // -1/6!
    pc = __builtin_fma(pc, t2, 4.1666666666666666667e-2);     // 1/4!
    pc = __builtin_fma(pc, t2, -0.5);                         // -1/2!
    const double ct = __builtin_fma(t2, pc, 1.0);
    const double2 e = tab[static_cast<int>(n) & 63];
    *s_out = __builtin_fma(e.x, ct, e.y * st);
    *c_out = __builtin_fma(e.y, ct, -(e.x * st));
}

// {sin, cos}(j pi/32), j < 64, correctly rounded (evaluated to 45 digits).
static __constant__ const double2 kSinCosPi32[64] = {
    {0.0, 1.0}, {0.0980171403295606, 0.9951847266721969},
    {0.19509032201612828, 0.9807852804032304}, {0.2902846772544624, 0.9569403357322088},
    {0.3826834323650898, 0.9238795325112867}, {0.47139673682599764, 0.881921264348355},
    {0.5555702330196022, 0.8314696123025452}, {0.6343932841636455, 0.773010453362737},
    {0.7071067811865476, 0.7071067811865476}, {0.773010453362737, 0.6343932841636455},
    {0.8314696123025452, 0.5555702330196022}, {0.881921264348355, 0.47139673682599764},
    {0.9238795325112867, 0.3826834323650898}, {0.9569403357322088, 0.2902846772544624},
    {0.9807852804032304, 0.19509032201612828}, {0.9951847266721969, 0.0980171403295606},
    {1.0, -1.6185598133277091e-46}, {0.9951847266721969, -0.0980171403295606},
    {0.9807852804032304, -0.19509032201612828}, {0.9569403357322088, -0.2902846772544624},
    {0.9238795325112867, -0.3826834323650898}, {0.881921264348355, -0.47139673682599764},
    {0.8314696123025452, -0.5555702330196022}, {0.773010453362737, -0.6343932841636455},
    {0.7071067811865476, -0.7071067811865476}, {0.6343932841636455, -0.773010453362737},
    {0.5555702330196022, -0.8314696123025452}, {0.47139673682599764, -0.881921264348355},
    {0.3826834323650898, -0.9238795325112867}, {0.2902846772544624, -0.9569403357322088},
    {0.19509032201612828, -0.9807852804032304}, {0.0980171403295606, -0.9951847266721969},
    {1.6936229514088527e-46, -1.0}, {-0.0980171403295606, -0.9951847266721969},
    {-0.19509032201612828, -0.9807852804032304}, {-0.2902846772544624, -0.9569403357322088},
    {-0.3826834323650898, -0.9238795325112867}, {-0.47139673682599764, -0.881921264348355},
    {-0.5555702330196022, -0.8314696123025452}, {-0.6343932841636455, -0.773010453362737},
    {-0.7071067811865476, -0.7071067811865476}, {-0.773010453362737, -0.6343932841636455},
    {-0.8314696123025452, -0.5555702330196022}, {-0.881921264348355, -0.47139673682599764},
    {-0.9238795325112867, -0.3826834323650898}, {-0.9569403357322088, -0.2902846772544624},
    {-0.9807852804032304, -0.19509032201612828}, {-0.9951847266721969, -0.0980171403295606},
    {-1.0, -9.059354245791371e-47}, {-0.9951847266721969, 0.0980171403295606},
    {-0.9807852804032304, 0.19509032201612828}, {-0.9569403357322088, 0.2902846772544624},
    {-0.9238795325112867, 0.3826834323650898}, {-0.881921264348355, 0.47139673682599764},
    {-0.8314696123025452, 0.5555702330196022}, {-0.773010453362737, 0.6343932841636455},
    {-0.7071067811865476, 0.7071067811865476}, {-0.6343932841636455, 0.773010453362737},
    {-0.5555702330196022, 0.8314696123025452}, {-0.47139673682599764, 0.881921264348355},
    {-0.3826834323650898, 0.9238795325112867}, {-0.2902846772544624, 0.9569403357322088},
    {-0.19509032201612828, 0.9807852804032304}, {-0.0980171403295606, 0.9951847266721969},
};

// Copies the table into LDS from the threads of a block (the caller
// synchronises before use).
__device__ __forceinline__ void sincos_table_fill(double2* tab) {
    const int j = threadIdx.x;
    if (j < 64) tab[j] = kSinCosPi32[j];
}

// 1/sqrt(x) in FP64: hardware estimate refined by one third-order step
// y <- y + y e (1/2 + 3/8 e), e = 1 - x y^2.  The estimate's relative error
// is below 2^-24 on gfx950 (measured over 1e-6 <= x <= 1e2), so the step
// leaves about 5/16 (2^-24)^3, far below one ulp; what remains is the
// rounding of the step itself.
__device__ __forceinline__ double rsqrt_f64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-x * y, y, 1.0);
    y = __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
    return y;
}

// FP32 helpers: hardware reciprocal square root and sin/cos (inputs in
// revolutions for the hardware units).
__device__ __forceinline__ float rsqrt_f32(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ void sincos_f32(float x, float* s, float* c) {
    const float inv2pi = 0.15915494309189533577f;
    float rev = x * inv2pi;
    rev = rev - __builtin_rintf(rev);
    *s = __builtin_amdgcn_sinf(rev);
    *c = __builtin_amdgcn_cosf(rev);
}

template <typename R> struct Math;
template <> struct Math<double> {
    static __device__ __forceinline__ double rsqrt(double x) { return rsqrt_f64(x); }
    static __device__ __forceinline__ void sincos(double x, double* s, double* c) {
        sincos_f64_tab(x, kSinCosPi32, s, c);
    }
};
template <> struct Math<float> {
    static __device__ __forceinline__ float rsqrt(float x) { return rsqrt_f32(x); }
    static __device__ __forceinline__ void sincos(float x, float* s, float* c) { sincos_f32(x, s, c); }
};

// P1 shape function of parametric slot `slot` at (r1, r2); slot < 0 is P0.
// src/managers/BEM_manager.py:404-425
template <typename R>
__device__ __forceinline__ R shape(R r1, R r2, int slot) {
    return slot == 0 ? R(1) - r1 : (slot == 1 ? r1 - r2 : (slot == 2 ? r2 : R(1)));
}

// Wave-level sum (64 lanes) via ds_swizzle-free shuffles.
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace bem
