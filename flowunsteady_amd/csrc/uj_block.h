// uj_block.h -- one workgroup's share of the all-pairs U and J: a block of
// consecutive targets against a contiguous range of sources, staged through
// LDS one tile of BLK sources at a time.  The single-field kernels call these
// with the whole field as the source range; the batched kernels call them with
// the range of the target's own field (fields never interact).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.h"

namespace vpmx {

// FP64 source record pair (x, y, z, Gx) (Gy, Gz, sigma, 1/sigma), as written by
// pack_f64_kernel.  Targets [t0, t1) read X (SoA, stride cap); sources
// [s0, s1).  Reference pair algebra: src/FLOWUnsteady_processing_force.jl:888-905
// extended to J = grad U.  `tile` holds 2*BLK records.  d = (U0..2, J0..8).
// DS: the core-size derivatives D[U], D[J] instead (g, g' -> h, h').
template <int K, int BLK, bool DS = false>
__device__ __forceinline__ void uj_f64_sums(const double4 *__restrict__ src, int64_t s0,
                                            int64_t s1, const double *__restrict__ X, int64_t t0,
                                            int64_t t1, int64_t cap, double4 *tile, double d[12])
{
    const int t = threadIdx.x;
    const int64_t i = t0 + t;
    const bool active = i < t1;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (active) {
        xi = X[i];
        yi = X[cap + i];
        zi = X[2 * cap + i];
    }
    double u0 = 0, u1 = 0, u2 = 0;
    double j0 = 0, j1 = 0, j2 = 0, j3 = 0, j4 = 0, j5 = 0, j6 = 0, j7 = 0, j8 = 0;

    for (int64_t base = s0; base < s1; base += BLK) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < s1) {
            tile[2 * t] = src[2 * js];
            tile[2 * t + 1] = src[2 * js + 1];
        }
        __syncthreads();
        const int cnt = (int)((s1 - base) < BLK ? (s1 - base) : BLK);
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 a = tile[2 * jj], b = tile[2 * jj + 1];
            const double dX1 = xi - a.x, dX2 = yi - a.y, dX3 = zi - a.z;
            const double r2 = dX1 * dX1 + dX2 * dX2 + dX3 * dX3;
            if (r2 > 0.0) {   // the reference skips r == 0
                const double r = sqrt(r2);
                double g, dg;
                if constexpr (DS)
                    h_dh_f64<K>(r * b.w, g, dg);
                else
                    g_dg_f64<K>(r * b.w, g, dg);
                const double rinv = 1.0 / r;
                const double r3i = rinv * rinv * rinv;
                const double G1 = a.w, G2 = b.x, G3 = b.y;
                const double c1 = -INV4PI * r3i * (dX2 * G3 - dX3 * G2);
                const double c2 = -INV4PI * r3i * (dX3 * G1 - dX1 * G3);
                const double c3 = -INV4PI * r3i * (dX1 * G2 - dX2 * G1);
                u0 += g * c1;
                u1 += g * c2;
                u2 += g * c3;
                const double aux = dg * b.w * rinv - 3.0 * g * rinv * rinv;
                j0 += aux * c1 * dX1; j1 += aux * c2 * dX1; j2 += aux * c3 * dX1;
                j3 += aux * c1 * dX2; j4 += aux * c2 * dX2; j5 += aux * c3 * dX2;
                j6 += aux * c1 * dX3; j7 += aux * c2 * dX3; j8 += aux * c3 * dX3;
                const double a2 = -INV4PI * g * r3i;
                j1 -= a2 * G3;
                j2 += a2 * G2;
                j3 += a2 * G3;
                j5 -= a2 * G1;
                j6 -= a2 * G2;
                j7 += a2 * G1;
            }
        }
    }
    d[0] = u0; d[1] = u1; d[2] = u2;
    d[3] = j0; d[4] = j1; d[5] = j2; d[6] = j3; d[7] = j4; d[8] = j5;
    d[9] = j6; d[10] = j7; d[11] = j8;
}

// ---------------------------------------------------------------------------
// FP64 on the Morton-sorted field.  Sources and targets are slots of the
// sorted records (x, y, z, Gx) (Gy, Gz, sigma, 1/sigma) (pack_f64_sorted);
// results go through perm.  The 15 sums are kept without the -1/(4 pi)
// factor: d = (u 0..2, w = sum g r^-3 Gamma 0..2, radial J 0..8 (k + 3 l)).
//
// Far field: for the gaussian kernels 1 - g(rho) < 3e-19 and rho g'(rho)
// < 2e-17 at rho >= 9.5, so g = 1 and g' = 0 to FP64 rounding there and the
// pair is the singular one: no erf / exp / sqrt / division.  A source tile
// whose gap to the target block is >= 9.5 sigma_max (> 0 for the singular
// kernel) runs f64_pair_far for every pair; within other tiles a wave whose
// every lane has rho^2 >= 90.25 for the current source takes it too.
constexpr double RHO2_FAR_F64 = 90.25;
constexpr double RHO_FAR_F64 = 9.5;

__device__ __forceinline__ void f64_pair_far(const double4 a, const double4 b, double xi,
                                             double yi, double zi, double d[15])
{
    const double dX1 = xi - a.x, dX2 = yi - a.y, dX3 = zi - a.z;
    const double r2 = dX1 * dX1 + dX2 * dX2 + dX3 * dX3;
    const double rinv = rsqrt(r2);
    const double ri2 = rinv * rinv, r3i = ri2 * rinv;
    const double G1 = a.w, G2 = b.x, G3 = b.y;
    const double c1 = r3i * (dX2 * G3 - dX3 * G2);
    const double c2 = r3i * (dX3 * G1 - dX1 * G3);
    const double c3 = r3i * (dX1 * G2 - dX2 * G1);
    d[0] += c1; d[1] += c2; d[2] += c3;
    d[3] += r3i * G1; d[4] += r3i * G2; d[5] += r3i * G3;
    const double aux = -3.0 * ri2;
    const double e1 = aux * c1, e2 = aux * c2, e3 = aux * c3;
    d[6] += e1 * dX1; d[7] += e2 * dX1; d[8] += e3 * dX1;
    d[9] += e1 * dX2; d[10] += e2 * dX2; d[11] += e3 * dX2;
    d[12] += e1 * dX3; d[13] += e2 * dX3; d[14] += e3 * dX3;
}

// Full evaluation (reference pair algebra, src/FLOWUnsteady_processing_force.jl:888-905
// extended to J); coincident pairs skipped.
template <int K>
__device__ __forceinline__ void f64_pair_full(const double4 a, const double4 b, double xi,
                                              double yi, double zi, double d[15])
{
    const double dX1 = xi - a.x, dX2 = yi - a.y, dX3 = zi - a.z;
    const double r2 = dX1 * dX1 + dX2 * dX2 + dX3 * dX3;
    if (!(r2 > 0.0)) return;   // the reference skips r == 0
    const double r = sqrt(r2);
    double g, dg;
    g_dg_f64<K>(r * b.w, g, dg);
    const double rinv = 1.0 / r;
    const double ri2 = rinv * rinv, r3i = ri2 * rinv;
    const double G1 = a.w, G2 = b.x, G3 = b.y;
    const double c1 = r3i * (dX2 * G3 - dX3 * G2);
    const double c2 = r3i * (dX3 * G1 - dX1 * G3);
    const double c3 = r3i * (dX1 * G2 - dX2 * G1);
    d[0] += g * c1; d[1] += g * c2; d[2] += g * c3;
    const double gr = g * r3i;
    d[3] += gr * G1; d[4] += gr * G2; d[5] += gr * G3;
    const double aux = dg * b.w * rinv - 3.0 * g * ri2;
    const double e1 = aux * c1, e2 = aux * c2, e3 = aux * c3;
    d[6] += e1 * dX1; d[7] += e2 * dX1; d[8] += e3 * dX1;
    d[9] += e1 * dX2; d[10] += e2 * dX2; d[11] += e3 * dX2;
    d[12] += e1 * dX3; d[13] += e2 * dX3; d[14] += e3 * dX3;
}

// One target block (slots [t0, t1), t0 a multiple of BLK) against source
// slots [s0, s1) (s0 a multiple of BLK).  tb: bounds of each BLK-slot tile
// (centre, radius), (largest sigma, -, -, -).  `tile` holds 2 BLK records.
template <int K, int BLK>
__device__ __forceinline__ void uj_f64s_sums(const double4 *__restrict__ src, int64_t s0,
                                             int64_t s1, int64_t t0, int64_t t1,
                                             const double4 *__restrict__ tb, double4 *tile,
                                             double d[15])
{
    const int t = threadIdx.x;
    const int64_t s = t0 + t;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    if (s < t1) {
        const double4 p = src[2 * s];
        xi = p.x;
        yi = p.y;
        zi = p.z;
    }
    const double4 T = tb[2 * (t0 / BLK)];
    for (int k = 0; k < 15; ++k) d[k] = 0.0;
    for (int64_t base = s0; base < s1; base += BLK) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < s1) {
            tile[2 * t] = src[2 * js];
            tile[2 * t + 1] = src[2 * js + 1];
        }
        __syncthreads();
        const int cnt = (int)((s1 - base) < BLK ? (s1 - base) : BLK);
        bool far = false;
        if constexpr (K == GAUSSIANERF || K == GAUSSIAN || K == SINGULAR) {
            if (cnt == BLK) {
                const int64_t ts = base / BLK;
                const double4 S = tb[2 * ts];
                const double ex = S.x - T.x, ey = S.y - T.y, ez = S.z - T.z;
                const double gap = sqrt(ex * ex + ey * ey + ez * ez) - S.w - T.w;
                if constexpr (K == SINGULAR)
                    far = gap > 0.0;
                else
                    far = gap >= RHO_FAR_F64 * tb[2 * ts + 1].x;
            }
        }
        if (far) {
#pragma unroll 2
            for (int jj = 0; jj < BLK; ++jj)
                f64_pair_far(tile[2 * jj], tile[2 * jj + 1], xi, yi, zi, d);
        } else {
            for (int jj = 0; jj < cnt; ++jj) {
                const double4 a = tile[2 * jj], b = tile[2 * jj + 1];
                bool lane_far = false;
                if constexpr (K == GAUSSIANERF || K == GAUSSIAN) {
                    const double dX1 = xi - a.x, dX2 = yi - a.y, dX3 = zi - a.z;
                    const double rho2 = (dX1 * dX1 + dX2 * dX2 + dX3 * dX3) * (b.w * b.w);
                    lane_far = rho2 >= RHO2_FAR_F64;
                }
                if (__all(lane_far))
                    f64_pair_far(a, b, xi, yi, zi, d);
                else
                    f64_pair_full<K>(a, b, xi, yi, zi, d);
            }
        }
    }
}

// U and J of particle i (stride ldo) from the 15 sums of uj_f64s_sums.
__device__ __forceinline__ void uj_f64s_store(const double d[15], int64_t i, int64_t ldo,
                                              double *__restrict__ U, double *__restrict__ J,
                                              int reset)
{
    const double k = -INV4PI;
    double v[12];
    v[0] = k * d[0]; v[1] = k * d[1]; v[2] = k * d[2];
    v[3] = k * d[6];
    v[4] = k * (d[7] - d[5]);
    v[5] = k * (d[8] + d[4]);
    v[6] = k * (d[9] + d[5]);
    v[7] = k * d[10];
    v[8] = k * (d[11] - d[3]);
    v[9] = k * (d[12] - d[4]);
    v[10] = k * (d[13] + d[3]);
    v[11] = k * d[14];
    for (int q = 0; q < 3; ++q) U[q * ldo + i] = reset ? v[q] : U[q * ldo + i] + v[q];
    for (int q = 0; q < 9; ++q) J[q * ldo + i] = reset ? v[3 + q] : J[q * ldo + i] + v[3 + q];
}

// U and J of target i (stride ldo) from the 12 FP64 sums.
__device__ __forceinline__ void uj_f64_store(const double d[12], int64_t i, int64_t ldo,
                                             double *__restrict__ U, double *__restrict__ J,
                                             int reset)
{
    for (int k = 0; k < 3; ++k) U[k * ldo + i] = reset ? d[k] : U[k * ldo + i] + d[k];
    for (int k = 0; k < 9; ++k) J[k * ldo + i] = reset ? d[3 + k] : J[k * ldo + i] + d[3 + k];
}

template <int K, int BLK, bool DS = false>
__device__ __forceinline__ void uj_f64_block(const double4 *__restrict__ src, int64_t s0,
                                             int64_t s1, const double *__restrict__ X,
                                             int64_t t0, int64_t t1, int64_t cap,
                                             double *__restrict__ U, double *__restrict__ J,
                                             int reset, double4 *tile)
{
    double d[12];
    uj_f64_sums<K, BLK, DS>(src, s0, s1, X, t0, t1, cap, tile, d);
    const int64_t i = t0 + threadIdx.x;
    if (i < t1) uj_f64_store(d, i, cap, U, J, reset);
}

// Mixed precision: the 14 FP64 sums of one target block against sources
// [s0, s1).  Targets are the sorted slots [t0, t1) of `tgt`; sources the
// sorted slots of (spos, sgam).  Coordinates are taken relative to the block's
// first target in FP64 and rounded to FP32; FP32 partials are folded into
// FP64 every MIXED_FOLD pairs (once per tile for far tiles).  d = (U0..2, W0..2, J0..7) before the -1/(4 pi) factor.
//
// With tile bounds `tb` (per BLK-slot tile counted from slot `torg`: (centre,
// radius), (largest sigma, -, -, -); the block's own targets are tile
// (t0 - torg)/BLK), a tile whose every pair is in the far field -- gap >=
// 7.001 sigma_max for the gaussian kernels, gap > 0 for the singular one --
// runs mixed_pair_far without per-pair tests, the same arithmetic the per-pair
// path takes for such pairs.  t0 - torg and s0 - torg must then be multiples
// of BLK.
//
// DS: the core-size derivatives D[U], D[J] in the same 14 sums
// (mixed_pair_dsig); far tiles are skipped, their pairs being beyond the
// cores (e^{-24.5} of the core value for the gaussian kernels).
template <int K, int BLK, bool DS = false>
__device__ __forceinline__ void uj_mixed_sums(const double4 *__restrict__ tgt, int64_t t0,
                                              int64_t t1, const double4 *__restrict__ spos,
                                              const float4 *__restrict__ sgam, int64_t s0,
                                              int64_t s1, const double4 *__restrict__ tb,
                                              int64_t torg, float4 *tpos, float4 *tgam,
                                              float4 *tlo, double d[14])
{
    static_assert(BLK % MIXED_FOLD == 0, "tiles hold whole fold groups");
    const int t = threadIdx.x;
    const int64_t s = t0 + t;
    const bool active = s < t1;
    const double4 o = tgt[t0];
    float xi = 0.f, yi = 0.f, zi = 0.f, xl = 0.f, yl = 0.f, zl = 0.f;
    if (active) {
        const double4 p = tgt[s];
        const double ox = p.x - o.x, oy = p.y - o.y, oz = p.z - o.z;
        xi = (float)ox;
        yi = (float)oy;
        zi = (float)oz;
        xl = (float)(ox - (double)xi);
        yl = (float)(oy - (double)yi);
        zl = (float)(oz - (double)zi);
    }
    double4 T = make_double4(0.0, 0.0, 0.0, 0.0);
    if (tb) T = tb[2 * ((t0 - torg) / BLK)];
    for (int k = 0; k < 14; ++k) d[k] = 0.0;

    for (int64_t base = s0; base < s1; base += BLK) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < s1) {
            const double4 p = spos[js];
            const double ox = p.x - o.x, oy = p.y - o.y, oz = p.z - o.z;
            const float hx = (float)ox, hy = (float)oy, hz = (float)oz;
            tpos[t] = make_float4(hx, hy, hz, (float)p.w);
            tlo[t] = make_float4((float)(ox - (double)hx), (float)(oy - (double)hy),
                                 (float)(oz - (double)hz), 0.f);
            tgam[t] = sgam[js];
        }
        __syncthreads();
        const int cnt = (int)((s1 - base) < BLK ? (s1 - base) : BLK);
        bool far = false;
        if constexpr (K == GAUSSIANERF || K == GAUSSIAN || K == SINGULAR) {
            if (tb && cnt == BLK) {
                const int64_t ts = (base - torg) / BLK;
                const double4 S = tb[2 * ts];
                const double ex = S.x - T.x, ey = S.y - T.y, ez = S.z - T.z;
                const double gap = sqrt(ex * ex + ey * ey + ez * ez) - S.w - T.w;
                if constexpr (K == SINGULAR)
                    far = gap > 0.0;
                else
                    far = gap >= 7.001 * tb[2 * ts + 1].x;
            }
        }
        MixedAcc acc;
        if constexpr (DS) {
            if (!far) {
                if (cnt == BLK) {
                    for (int jb = 0; jb < BLK; jb += MIXED_FOLD) {
#pragma unroll 4
                        for (int jj = jb; jj < jb + MIXED_FOLD; ++jj)
                            mixed_pair_dsig_hl<K>(tpos[jj], tlo[jj], tgam[jj], xi, yi, zi, xl,
                                                  yl, zl, acc);
                        mixed_fold(d, acc);
                    }
                } else {
                    for (int jj = 0; jj < cnt; ++jj) {
                        mixed_pair_dsig_hl<K>(tpos[jj], tlo[jj], tgam[jj], xi, yi, zi, xl, yl,
                                              zl, acc);
                        if ((jj & (MIXED_FOLD - 1)) == MIXED_FOLD - 1) mixed_fold(d, acc);
                    }
                }
            }
        } else if (far) {
            // far pairs are at least 7 core sizes (or a gap) away: their terms
            // stay small and one partial per tile suffices
#pragma unroll 4
            for (int jj = 0; jj < BLK; ++jj)
                mixed_pair_far<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
        } else if (cnt == BLK) {
            for (int jb = 0; jb < BLK; jb += MIXED_FOLD) {
#pragma unroll 4
                for (int jj = jb; jj < jb + MIXED_FOLD; ++jj)
                    mixed_pair_hl<K>(tpos[jj], &tlo[jj], tgam[jj], xi, yi, zi, xl, yl, zl, acc);
                mixed_fold(d, acc);
            }
        } else {
            for (int jj = 0; jj < cnt; ++jj) {
                mixed_pair_hl<K>(tpos[jj], &tlo[jj], tgam[jj], xi, yi, zi, xl, yl, zl, acc);
                if ((jj & (MIXED_FOLD - 1)) == MIXED_FOLD - 1) mixed_fold(d, acc);
            }
        }
        mixed_fold(d, acc);
    }
}

// U and J of target i from the 14 sums: J = -1/(4 pi) (radial + eps_klm W_m),
// J22 from the zero trace; reset = 0 adds to what U, J hold.
__device__ __forceinline__ void uj_mixed_store(const double d[14], int64_t i, int64_t ldo,
                                               double *__restrict__ U, double *__restrict__ J,
                                               int reset)
{
    const double k = -INV4PI;
    double u0 = k * d[0], u1 = k * d[1], u2 = k * d[2];
    double j0 = k * d[6], j1 = k * (d[7] - d[5]), j2 = k * (d[8] + d[4]);
    double j3 = k * (d[9] + d[5]), j4 = k * d[10], j5 = k * (d[11] - d[3]);
    double j6 = k * (d[12] - d[4]), j7 = k * (d[13] + d[3]), j8 = -k * (d[6] + d[10]);
    if (!reset) {
        u0 += U[i]; u1 += U[ldo + i]; u2 += U[2 * ldo + i];
        j0 += J[i]; j1 += J[ldo + i]; j2 += J[2 * ldo + i];
        j3 += J[3 * ldo + i]; j4 += J[4 * ldo + i]; j5 += J[5 * ldo + i];
        j6 += J[6 * ldo + i]; j7 += J[7 * ldo + i]; j8 += J[8 * ldo + i];
    }
    U[i] = u0; U[ldo + i] = u1; U[2 * ldo + i] = u2;
    J[i] = j0; J[ldo + i] = j1; J[2 * ldo + i] = j2;
    J[3 * ldo + i] = j3; J[4 * ldo + i] = j4; J[5 * ldo + i] = j5;
    J[6 * ldo + i] = j6; J[7 * ldo + i] = j7; J[8 * ldo + i] = j8;
}

// One block: sums over [s0, s1) and the store through perm (stride ldo).
template <int K, int BLK, bool DS = false>
__device__ __forceinline__ void uj_mixed_block(const double4 *__restrict__ tgt, int64_t t0,
                                               int64_t t1, const int32_t *__restrict__ perm,
                                               const double4 *__restrict__ spos,
                                               const float4 *__restrict__ sgam, int64_t s0,
                                               int64_t s1, int64_t ldo,
                                               double *__restrict__ U, double *__restrict__ J,
                                               int reset, float4 *tpos, float4 *tgam,
                                               float4 *tlo,
                                               const double4 *__restrict__ tb = nullptr,
                                               int64_t torg = 0)
{
    double d[14];
    uj_mixed_sums<K, BLK, DS>(tgt, t0, t1, spos, sgam, s0, s1, tb, torg, tpos, tgam, tlo, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < t1) uj_mixed_store(d, perm[s], ldo, U, J, reset);
}

}  // namespace vpmx
