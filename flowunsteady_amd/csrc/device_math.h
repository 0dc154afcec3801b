// device_math.h -- radial kernel functions of the regularised Biot-Savart law.
//
// Kernels (docs/src/theory/rvpm.md:84,89-99,365,380), rho = r/sigma:
//   gaussianerf  zeta = (2 pi)^-3/2 e^{-rho^2/2}
//                g    = erf(rho/sqrt2) - sqrt(2/pi) rho e^{-rho^2/2}
//                g'   = sqrt(2/pi) rho^2 e^{-rho^2/2}
//   winckelmans  zeta = 15/(8 pi) (rho^2+1)^{-7/2}
//                g    = rho^3 (rho^2 + 5/2) (rho^2+1)^{-5/2}
//                g'   = 15/2 rho^2 (rho^2+1)^{-7/2}
//   gaussian     zeta = 3/(4 pi) e^{-rho^3},  g = 1 - e^{-rho^3},  g' = 3 rho^2 e^{-rho^3}
//   singular     g = 1, g' = 0
//
// A pair (target i, source j, dX = x_i - x_j, r = |dX|) contributes
//   U_k    += -1/(4 pi) A c_k                        c = dX x Gamma_j
//   J[k,l] += -1/(4 pi) ( B c_k dX_l + A eps_klm Gamma_j,m )
// with the two radial coefficients
//   A = g / r^3,     B = g'/(sigma r^4) - 3 g / r^5,
// the same algebra as src/FLOWUnsteady_processing_force.jl:895-905 extended to
// J = grad U.  Both coefficients are finite as r -> 0 (A -> g/rho^3 sigma^-3).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vpmx {

enum Kernel : int { SINGULAR = 0, GAUSSIAN = 1, GAUSSIANERF = 2, WINCKELMANS = 3 };

constexpr double INV4PI = 0.07957747154594767;   // 1/(4 pi), processing_force.jl:900
constexpr double SQRT2_PI = 0.7978845608028654;  // sqrt(2/pi)
constexpr double INV_SQRT2 = 0.7071067811865476;

// ---------------------------------------------------------------------------
// FP64, reference form: g and g' exactly as the closed forms above (libm-style
// erf/exp from the ROCm device math library), A and B assembled as in the
// reference pair loop.  This is the correctness path.
template <int K>
__device__ __forceinline__ void g_dg_f64(double rho, double &g, double &dg)
{
    if constexpr (K == GAUSSIANERF) {
        double e = exp(-0.5 * rho * rho);
        g = erf(rho * INV_SQRT2) - SQRT2_PI * rho * e;
        dg = SQRT2_PI * rho * rho * e;
    } else if constexpr (K == GAUSSIAN) {
        double e = exp(-rho * rho * rho);
        g = 1.0 - e;
        dg = 3.0 * rho * rho * e;
    } else if constexpr (K == WINCKELMANS) {
        double q = rho * rho + 1.0;
        double sq = sqrt(q);
        double q52 = q * q * sq;
        g = rho * rho * rho * (rho * rho + 2.5) / q52;
        dg = 7.5 * rho * rho / (q52 * q);
    } else {
        g = 1.0;
        dg = 0.0;
    }
}

template <int K>
__device__ __forceinline__ double zeta_f64(double rho)
{
    if constexpr (K == GAUSSIANERF) {
        return 0.06349363593424097 * exp(-0.5 * rho * rho);   // (2 pi)^-3/2
    } else if constexpr (K == GAUSSIAN) {
        return 0.238732414637843 * exp(-rho * rho * rho);     // 3/(4 pi)
    } else if constexpr (K == WINCKELMANS) {
        double q = rho * rho + 1.0;
        double sq = sqrt(q);
        return 0.5968310365946075 / (q * q * q * sq);          // 15/(8 pi)
    } else {
        return 0.0;
    }
}

// eta(rho) = -zeta'(rho)/rho, the PSE kernel (second moment 1 for every
// smooth kernel).
template <int K>
__device__ __forceinline__ double eta_f64(double rho)
{
    const double z = zeta_f64<K>(rho);
    if constexpr (K == GAUSSIANERF) return z;
    else if constexpr (K == GAUSSIAN) return 3.0 * rho * z;
    else if constexpr (K == WINCKELMANS) return 7.0 * z / (rho * rho + 1.0);
    else return 0.0;
}

// ---------------------------------------------------------------------------
// Core-size derivatives for the dynamic SFS procedure (docs/src/theory/rvpm.md:276-290),
// every core scaled together (D = sigma d/dsigma):
//   h  = D g(r/sigma)  = -rho g'(rho) = -4 pi rho^3 zeta(rho)
//   h' = dh/drho       = -4 pi (3 rho^2 zeta + rho^3 zeta')
//   zD = sigma^3 D [zeta(r/sigma)/sigma^3] = -(3 zeta + rho zeta')
// U and J carry g only through g(r/sigma)/r^3, so D[U] and D[J] are the pair
// algebra of U and J with (g, g') replaced by (h, h'); in radial-coefficient
// form A_D = h/r^3 = -4 pi zeta/sigma^3 and B_D = -4 pi zeta'/(rho sigma^5),
// both finite at r = 0 and negligible beyond the core (no far field).
template <int K>
__device__ __forceinline__ void h_dh_f64(double rho, double &h, double &dh)
{
    if constexpr (K == GAUSSIANERF) {
        const double e = exp(-0.5 * rho * rho);
        const double r2 = rho * rho;
        h = -SQRT2_PI * r2 * rho * e;
        dh = -SQRT2_PI * r2 * (3.0 - r2) * e;
    } else if constexpr (K == GAUSSIAN) {
        const double r3 = rho * rho * rho;
        const double e = exp(-r3);
        h = -3.0 * r3 * e;
        dh = -9.0 * rho * rho * (1.0 - r3) * e;
    } else if constexpr (K == WINCKELMANS) {
        const double r2 = rho * rho, q = r2 + 1.0;
        const double q72 = q * q * q * sqrt(q);
        h = -7.5 * r2 * rho / q72;
        dh = -7.5 * (3.0 * r2 - 7.0 * r2 * r2 / q) / q72;
    } else {
        h = 0.0;
        dh = 0.0;
    }
}

// zeta(rho) and zD(rho) together (one exp).
template <int K>
__device__ __forceinline__ void zeta_zD_f64(double rho, double &z, double &zD)
{
    if constexpr (K == GAUSSIANERF) {
        z = 0.06349363593424097 * exp(-0.5 * rho * rho);
        zD = (rho * rho - 3.0) * z;
    } else if constexpr (K == GAUSSIAN) {
        const double r3 = rho * rho * rho;
        z = 0.238732414637843 * exp(-r3);
        zD = 3.0 * (r3 - 1.0) * z;
    } else if constexpr (K == WINCKELMANS) {
        const double q = rho * rho + 1.0;
        z = 0.5968310365946075 / (q * q * q * sqrt(q));
        zD = (4.0 - 7.0 / q) * z;
    } else {
        z = 0.0;
        zD = 0.0;
    }
}

// ---------------------------------------------------------------------------
// FP32 forms used by the mixed-precision kernels.
//
// gaussianerf, 1.5 <= rho < 7: erfc(x) = t P(t) e^{-x^2}, t = 1/(1 + p x),
//   x = rho/sqrt2, P of degree 4 -- the form of Hastings' approximation
//   (Abramowitz & Stegun 7.1.26), with p and the five coefficients refitted
//   here as a weighted minimax approximation of erfc(x) e^{x^2} on exactly
//   this range, weighted by e^{-x^2} / |rho g' - 3 g| (the factor by which an
//   error in erfc reaches B).  Its e^{-x^2} is the same e^{-rho^2/2} that g'
//   needs, so one hardware exp serves both: g = 1 - e (t P(t) + sqrt(2/pi) rho).
//   Evaluated in FP32 it puts at most 2.0e-7 relative error into B and 8e-8
//   into A (A&S 7.1.26 on the same range: 5.5e-7 and 2.1e-7).
// gaussianerf, rho < 1.5: Taylor series of g/rho^3 and of
//   sigma^5 B = (g'/rho^2 - 3 g/rho^3)/rho^2 obtained by integrating the series
//   of g' = sqrt(2/pi) rho^2 e^{-rho^2/2} term by term:
//   g/rho^3      = sqrt(2/pi) sum_n (-1)^n rho^{2n} / (2^n n! (2n+3))
//   sigma^5 B    = sqrt(2/pi) sum_{n>=1} (-1)^n 2n rho^{2n-2} / (2^n n! (2n+3))
//   which avoid the cancellation of the closed form as rho -> 0.
// rho^2 >= 49 for every lane of a wave: g = 1 and g' = 0 to FP32 rounding
//   (e^{-24.5} rho^3 < 1e-8), so A = 1/r^3, B = -3/r^5 without exp.
// winckelmans (any rho): with s^2 = r^2 + sigma^2,
//   A = (r^2 + 5/2 sigma^2) s^-5,   B = -3 (r^2 + 7/2 sigma^2) s^-7,
//   an exact rewrite of the closed forms that needs one rsqrt.

constexpr float RHO2_FAR = 49.0f;      // far-field fast path threshold (rho >= 7)
constexpr float RHO2_SERIES = 2.25f;   // series below rho = 1.5
constexpr float RHO3_SERIES_G = 0.5f;  // gaussian kernel: series below rho^3 = 0.5

// Hardware transcendentals on the GPU (v_exp_f32, v_rcp_f32, v_rsq_f32, each
// within 1 ulp); the host build of these formulas (used only by the CPU unit
// test of their accuracy) takes the C library equivalents.
#define VPMX_HD __host__ __device__ __forceinline__
VPMX_HD float fexp2(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_exp2f(x);
#else
    return exp2f(x);
#endif
}
VPMX_HD float frcp(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
VPMX_HD float frsq(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}

// Near the core (x = rho^2 < 2.25) g/rho^3 and sigma^5 B are the series
//   g/rho^3   = sum_n cg(n) x^n,          cg(n) = sqrt(2/pi) (-1)^n / (2^n n! (2n+3))
//   sigma^5 B = sum_{n>=1} cb(n) x^(n-1), cb(n) = 2 n cg(n)
// evaluated here as polynomials in x of degree 6 (A) and 7 (B): relative
// minimax fits of the series on [0, 2.25] (weighted least squares on 400
// Chebyshev nodes with Lawson reweighting, coefficients then rounded to
// FP32).  By FP32 Horner their largest relative errors over the range are
// 1.5e-7 (A) and 2.1e-7 (B), where the twelve-term Taylor sums they replace
// reach 2.1e-7 and 2.5e-7 (the FP32 rounding of an alternating sum sets
// both), in 13 instead of 22 operations.
constexpr float FIT_A[7] = {2.659615278e-01f, -7.978837192e-02f, 1.424734108e-02f,
                            -1.845401130e-03f, 1.869333064e-04f, -1.468154551e-05f,
                            7.031323435e-07f};
constexpr float FIT_B[8] = {-1.595769078e-01f, 5.699175224e-02f, -1.108170021e-02f,
                            1.511041657e-03f, -1.596564252e-04f, 1.368616176e-05f,
                            -9.297319252e-07f, 3.905145363e-08f};

VPMX_HD float series_hA(float x)  // g/rho^3
{
    float h = FIT_A[6];
#pragma unroll
    for (int n = 5; n >= 0; --n) h = fmaf(h, x, FIT_A[n]);
    return h;
}

VPMX_HD float series_hB(float x)  // sigma^5 B
{
    float h = FIT_B[7];
#pragma unroll
    for (int n = 6; n >= 0; --n) h = fmaf(h, x, FIT_B[n]);
    return h;
}

// Gaussian kernel (zeta = 3/(4 pi) e^{-rho^3}), u = rho^3 < 1/2: from the series
// of e^{-u},
//   g/rho^3   = sum_{n>=0} (-1)^n u^n / (n+1)!
//   sigma^5 B = 3 rho sum_{n>=1} (-1)^n n u^{n-1} / (n+1)!
constexpr int NSERIES_G = 9;
constexpr double fact(int n)
{
    double d = 1.0;
    for (int k = 2; k <= n; ++k) d *= k;
    return d;
}
constexpr float gg(int n) { return (float)((n % 2 ? -1.0 : 1.0) / fact(n + 1)); }
constexpr float gb(int n) { return (float)((n % 2 ? -1.0 : 1.0) * 3.0 * n / fact(n + 1)); }

VPMX_HD float series_gauss_hA(float u)
{
    float h = gg(NSERIES_G - 1);
#pragma unroll
    for (int n = NSERIES_G - 2; n >= 0; --n) h = fmaf(h, u, gg(n));
    return h;
}
VPMX_HD float series_gauss_hB(float u, float rho)
{
    float h = gb(NSERIES_G);
#pragma unroll
    for (int n = NSERIES_G - 1; n >= 1; --n) h = fmaf(h, u, gb(n));
    return h * rho;
}

// Constants of the erfc fit above (p, then the coefficients of t^1 .. t^5).
constexpr float ERFC_P = 0.5199999809265137f;
constexpr float ERFC_C1 = 0.2891683578491211f, ERFC_C2 = 0.33505475521087646f,
                ERFC_C3 = 0.08860228955745697f, ERFC_C4 = 0.495405912399292f,
                ERFC_C5 = -0.20866064727306366f;
constexpr float LOG2E = 1.4426950408889634f;

// Radial coefficients (A, B) of one pair in FP32, full evaluation (near field).
// r2 > 0 is guaranteed by the caller's mask (rinv = 0 for coincident pairs);
// p0, p1 are the per-source constants: gaussian(-erf) p0 = 1/sigma,
// p1 = 1/sigma^2; winckelmans p0 = sigma^2.
template <int K>
VPMX_HD void radial_near_f32(float r2, float rinv, float p0, float p1,
                                                float &A, float &B)
{
    if constexpr (K == GAUSSIANERF) {
        float rho2 = r2 * p1;
        if (rho2 < RHO2_SERIES) {
            float is3 = p1 * p0, is5 = is3 * p1;
            A = series_hA(rho2) * is3;
            B = series_hB(rho2) * is5;
        } else {
            float rinv2 = rinv * rinv, rinv3 = rinv2 * rinv;
            float rho = r2 * rinv * p0;
            float e = fexp2(-0.5f * LOG2E * rho2);
            float t = frcp(fmaf(ERFC_P * (float)INV_SQRT2, rho, 1.0f));
            float P = fmaf(fmaf(fmaf(fmaf(ERFC_C5, t, ERFC_C4), t, ERFC_C3), t, ERFC_C2), t, ERFC_C1) * t;
            float g = fmaf(-e, fmaf((float)SQRT2_PI, rho, P), 1.0f);
            float dg = (float)SQRT2_PI * rho2 * e;
            A = g * rinv3;
            B = (dg * p0 * rinv - 3.0f * g * rinv2) * rinv3;
        }
    } else if constexpr (K == GAUSSIAN) {
        float rho = r2 * rinv * p0;
        float u = rho * rho * rho;
        if (u < RHO3_SERIES_G) {
            float is3 = p1 * p0, is5 = is3 * p1;
            A = series_gauss_hA(u) * is3;
            B = series_gauss_hB(u, rho) * is5;
        } else {
            float rinv2 = rinv * rinv, rinv3 = rinv2 * rinv;
            float e = fexp2(-LOG2E * u);
            float g = 1.0f - e;
            float dg = 3.0f * rho * rho * e;
            A = g * rinv3;
            B = (dg * p0 * rinv - 3.0f * g * rinv2) * rinv3;
        }
    } else if constexpr (K == WINCKELMANS) {
        float s2 = r2 + p0;
        float si = frsq(s2);
        float si2 = si * si;
        float si5 = si2 * si2 * si;
        A = fmaf(2.5f, p0, r2) * si5;
        B = -3.0f * fmaf(3.5f, p0, r2) * si5 * si2;
    } else {
        float rinv2 = rinv * rinv;
        A = rinv2 * rinv;
        B = -3.0f * A * rinv2;
    }
}

// Radial coefficients (A_D, B_D) of the core-size derivative of one pair in
// FP32 (see h_dh_f64); p0, p1 as for radial_near_f32.
//   gaussianerf  A_D = -sqrt(2/pi) e^{-rho^2/2} / sigma^3,  B_D = sqrt(2/pi) e^{-rho^2/2} / sigma^5
//   gaussian     A_D = -3 e^{-rho^3} / sigma^3,             B_D = 9 rho e^{-rho^3} / sigma^5
//   winckelmans  A_D = -15/2 sigma^4 s^-7,                 B_D = 105/2 sigma^4 s^-9,  s^2 = r^2 + sigma^2
template <int K>
VPMX_HD void radial_dsig_f32(float r2, float p0, float p1, float &A, float &B)
{
    if constexpr (K == GAUSSIANERF) {
        const float is3 = p0 * p1;
        const float e = (float)SQRT2_PI * fexp2(-0.5f * LOG2E * r2 * p1);
        A = -e * is3;
        B = e * is3 * p1;
    } else if constexpr (K == GAUSSIAN) {
        const float is3 = p0 * p1;
        const float rho = sqrtf(r2 * p1);
        const float e = fexp2(-LOG2E * rho * rho * rho);
        A = -3.0f * e * is3;
        B = 9.0f * rho * e * is3 * p1;
    } else if constexpr (K == WINCKELMANS) {
        const float si = frsq(r2 + p0);
        const float si2 = si * si;
        const float si7 = si2 * si2 * si2 * si;
        const float s4 = p0 * p0;
        A = -7.5f * s4 * si7;
        B = 52.5f * s4 * si7 * si2;
    } else {
        A = 0.f;
        B = 0.f;
    }
}

// FP32 tile partials of one target: U, W = sum A Gamma (Kronecker part of J)
// and eight radial entries of J (the ninth follows from trace J = 0).
struct MixedAcc {
    float U0 = 0.f, U1 = 0.f, U2 = 0.f, W0 = 0.f, W1 = 0.f, W2 = 0.f;
    float J0 = 0.f, J1 = 0.f, J2 = 0.f, J3 = 0.f, J4 = 0.f, J5 = 0.f, J6 = 0.f, J7 = 0.f;
};

// Pairs one MixedAcc sums before it is folded into FP64.  A target's nearest
// sources give terms of order |Gamma|/sigma^3 that largely cancel between W
// and the radial part of J, and every later FP32 add onto such a partial
// rounds at 2^-24 of it: with the sources in Morton order those neighbours
// share a tile, and 256-pair partials lost up to 1e-5 of J on a ring of
// 8946 particles (sigma = 0.027), 32-pair partials 2e-7.
constexpr int MIXED_FOLD = 32;

// d[0..13] += the partials (U, W, J0..7), which restart from zero.
__device__ __forceinline__ void mixed_fold(double d[14], MixedAcc &acc)
{
    d[0] += acc.U0; d[1] += acc.U1; d[2] += acc.U2;
    d[3] += acc.W0; d[4] += acc.W1; d[5] += acc.W2;
    d[6] += acc.J0; d[7] += acc.J1; d[8] += acc.J2; d[9] += acc.J3;
    d[10] += acc.J4; d[11] += acc.J5; d[12] += acc.J6; d[13] += acc.J7;
    acc = MixedAcc();
}

// Accumulates one pair with radial coefficients A = g/r^3 and
// B = g'/(sigma r^4) - 3 g / r^5 (U, W = sum A Gamma, radial part of J).
__device__ __forceinline__ void mixed_accumulate(float dx, float dy, float dz, float A, float B,
                                                 const float4 b, MixedAcc &acc);

// One pair in FP32.  a = (source position relative to the block origin, p0),
// b = (Gamma, p1).  When every lane of the wave is in the far field
// (rho^2 >= 49) the gaussian kernels skip exp/erf: g = 1, g' = 0 to FP32
// rounding.  Coincident pairs contribute nothing.
template <int K>
__device__ __forceinline__ void mixed_pair_d(float dx, float dy, float dz, float p0,
                                             const float4 b, MixedAcc &acc);

template <int K>
__device__ __forceinline__ void mixed_pair(const float4 a, const float4 b, float xi, float yi,
                                           float zi, MixedAcc &acc)
{
    mixed_pair_d<K>(xi - a.x, yi - a.y, zi - a.z, a.w, b, acc);
}

// The same pair from its separation dX (target minus source) in FP32.
template <int K>
__device__ __forceinline__ void mixed_pair_d(float dx, float dy, float dz, float p0,
                                             const float4 b, MixedAcc &acc)
{
    const float4 a = make_float4(0.f, 0.f, 0.f, p0);
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool self = !(r2 > 0.f);
    const float rinv = self ? 0.f : frsq(r2);
    float A, B;
    if constexpr (K == GAUSSIANERF || K == GAUSSIAN) {
        const bool near = r2 * b.w < RHO2_FAR;
        if (__any(near)) {
            radial_near_f32<K>(r2, rinv, a.w, b.w, A, B);
        } else {
            const float ri2 = rinv * rinv;
            A = ri2 * rinv;
            B = -3.0f * A * ri2;
        }
    } else {
        radial_near_f32<K>(r2, rinv, a.w, b.w, A, B);
    }
    A = self ? 0.f : A;
    B = self ? 0.f : B;
    mixed_accumulate(dx, dy, dz, A, B, b, acc);
}

// The core-size derivative of the same pair (D[U], D[J] in the U, J slots of
// the accumulator).  Coincident pairs contribute nothing.
template <int K>
__device__ __forceinline__ void mixed_pair_dsig_d(float dx, float dy, float dz, float p0,
                                                  const float4 b, MixedAcc &acc);
template <int K>
__device__ __forceinline__ void mixed_pair_dsig(const float4 a, const float4 b, float xi,
                                                float yi, float zi, MixedAcc &acc)
{
    const float dx = xi - a.x, dy = yi - a.y, dz = zi - a.z;
    mixed_pair_dsig_d<K>(dx, dy, dz, a.w, b, acc);
}

// The same with (hi, lo) offsets of both ends: separation = hi difference +
// lo difference (these pairs all lie within a few cores).
template <int K>
__device__ __forceinline__ void mixed_pair_dsig_hl(const float4 a, const float4 l, const float4 b,
                                                   float xi, float yi, float zi, float xl,
                                                   float yl, float zl, MixedAcc &acc)
{
    const float dx = (xi - a.x) + (xl - l.x), dy = (yi - a.y) + (yl - l.y),
                dz = (zi - a.z) + (zl - l.z);
    mixed_pair_dsig_d<K>(dx, dy, dz, a.w, b, acc);
}

template <int K>
__device__ __forceinline__ void mixed_pair_dsig_d(float dx, float dy, float dz, float p0,
                                                  const float4 b, MixedAcc &acc)
{
    const float4 a = make_float4(0.f, 0.f, 0.f, p0);
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const bool self = !(r2 > 0.f);
    float A, B;
    radial_dsig_f32<K>(r2, a.w, b.w, A, B);
    A = self ? 0.f : A;
    B = self ? 0.f : B;
    mixed_accumulate(dx, dy, dz, A, B, b, acc);
}

// One pair with (hi, lo) FP32 offsets of both ends from the block origin:
// hi = (float)(x - o), lo = (float)((x - o) - hi).  While every lane of the
// wave has rho^2 >= 49 by the hi separation, the pair takes the far branch on
// the hi separation (its relative rounding is small against r >= 7 sigma);
// otherwise the whole wave takes the near branch on the separation hi
// difference + lo difference, which keeps close pairs at FP32 relative
// accuracy however far the block origin lies.  The lo record `alo` is read
// only there.  p1 = 1/sigma^2 for every kernel.  Coincident pairs (zero
// separation) contribute nothing.
template <int K>
__device__ __forceinline__ void mixed_pair_hl(const float4 a, const float4 *alo, const float4 b,
                                              float xi, float yi, float zi, float xl, float yl,
                                              float zl, MixedAcc &acc)
{
    float dx = xi - a.x, dy = yi - a.y, dz = zi - a.z;
    float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    float A, B;
    if (__any(r2 * b.w < RHO2_FAR)) {
        const float4 l = *alo;
        dx += xl - l.x;
        dy += yl - l.y;
        dz += zl - l.z;
        r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const bool self = !(r2 > 0.f);
        const float rinv = self ? 0.f : frsq(r2);
        radial_near_f32<K>(r2, rinv, a.w, b.w, A, B);
        A = self ? 0.f : A;
        B = self ? 0.f : B;
    } else if constexpr (K == WINCKELMANS) {
        radial_near_f32<K>(r2, 0.f, a.w, b.w, A, B);   // no far-field form: full evaluation
    } else {
        const float rinv = frsq(r2);                   // r2 > 0: rho^2 >= 49
        const float ri2 = rinv * rinv;
        A = ri2 * rinv;
        B = -3.0f * A * ri2;
    }
    mixed_accumulate(dx, dy, dz, A, B, b, acc);
}

// A pair known to lie in the far field: rho^2 >= 49 for the gaussian kernels
// (g = 1, g' = 0 to FP32 rounding), r > 0 for the singular kernel.  The same
// operations as the far branch of mixed_pair_d without the per-pair tests.
template <int K>
__device__ __forceinline__ void mixed_pair_far(const float4 a, const float4 b, float xi,
                                               float yi, float zi, MixedAcc &acc)
{
    const float dx = xi - a.x, dy = yi - a.y, dz = zi - a.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float rinv = frsq(r2);
    const float ri2 = rinv * rinv;
    const float A = ri2 * rinv;
    const float B = -3.0f * A * ri2;
    mixed_accumulate(dx, dy, dz, A, B, b, acc);
}

__device__ __forceinline__ void mixed_accumulate(float dx, float dy, float dz, float A, float B,
                                                 const float4 b, MixedAcc &acc)
{
    // c = dX x Gamma
    const float c0 = fmaf(dy, b.z, -dz * b.y);
    const float c1 = fmaf(dz, b.x, -dx * b.z);
    const float c2 = fmaf(dx, b.y, -dy * b.x);
    acc.U0 = fmaf(A, c0, acc.U0);
    acc.U1 = fmaf(A, c1, acc.U1);
    acc.U2 = fmaf(A, c2, acc.U2);
    acc.W0 = fmaf(A, b.x, acc.W0);
    acc.W1 = fmaf(A, b.y, acc.W1);
    acc.W2 = fmaf(A, b.z, acc.W2);
    const float ex = B * dx, ey = B * dy, ez = B * dz;
    // radial part c_k dX_l B, column-major (k + 3 l)
    acc.J0 = fmaf(c0, ex, acc.J0);
    acc.J1 = fmaf(c1, ex, acc.J1);
    acc.J2 = fmaf(c2, ex, acc.J2);
    acc.J3 = fmaf(c0, ey, acc.J3);
    acc.J4 = fmaf(c1, ey, acc.J4);
    acc.J5 = fmaf(c2, ey, acc.J5);
    acc.J6 = fmaf(c0, ez, acc.J6);
    acc.J7 = fmaf(c1, ez, acc.J7);
}

// 30-bit Morton key of a point in the cube of side max(extent) anchored at the
// box's low corner (box = lo xyz, hi xyz): 10 bits per axis, x in bit 0.
__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton30(const double *box, double x, double y, double z)
{
    const double ext = fmax(fmax(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
    const double scale = ext > 0.0 ? 1023.999 / ext : 0.0;
    const double p[3] = {x, y, z};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        double v = (p[k] - box[k]) * scale;
        v = fmin(fmax(v, 0.0), 1023.0);
        q[k] = (uint32_t)v;
    }
    return spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
}

}  // namespace vpmx
