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
// first target in FP64 and rounded to FP32; each tile's FP32 partials are
// folded into FP64.  d = (U0..2, W0..2, J0..7) before the -1/(4 pi) factor.
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
                                              double d[14])
{
    const int t = threadIdx.x;
    const int64_t s = t0 + t;
    const bool active = s < t1;
    const double4 o = tgt[t0];
    float xi = 0.f, yi = 0.f, zi = 0.f;
    if (active) {
        const double4 p = tgt[s];
        xi = (float)(p.x - o.x);
        yi = (float)(p.y - o.y);
        zi = (float)(p.z - o.z);
    }
    double4 T = make_double4(0.0, 0.0, 0.0, 0.0);
    if (tb) T = tb[2 * ((t0 - torg) / BLK)];
    for (int k = 0; k < 14; ++k) d[k] = 0.0;

    for (int64_t base = s0; base < s1; base += BLK) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < s1) {
            const double4 p = spos[js];
            tpos[t] = make_float4((float)(p.x - o.x), (float)(p.y - o.y), (float)(p.z - o.z),
                                  (float)p.w);
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
#pragma unroll 4
                    for (int jj = 0; jj < BLK; ++jj)
                        mixed_pair_dsig<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
                } else {
                    for (int jj = 0; jj < cnt; ++jj)
                        mixed_pair_dsig<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
                }
            }
        } else if (far) {
#pragma unroll 4
            for (int jj = 0; jj < BLK; ++jj)
                mixed_pair_far<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
        } else if (cnt == BLK) {
#pragma unroll 4
            for (int jj = 0; jj < BLK; ++jj)
                mixed_pair<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
        } else {
            for (int jj = 0; jj < cnt; ++jj)
                mixed_pair<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
        }
        d[0] += acc.U0; d[1] += acc.U1; d[2] += acc.U2;
        d[3] += acc.W0; d[4] += acc.W1; d[5] += acc.W2;
        d[6] += acc.J0; d[7] += acc.J1; d[8] += acc.J2; d[9] += acc.J3;
        d[10] += acc.J4; d[11] += acc.J5; d[12] += acc.J6; d[13] += acc.J7;
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
                                               const double4 *__restrict__ tb = nullptr,
                                               int64_t torg = 0)
{
    double d[14];
    uj_mixed_sums<K, BLK, DS>(tgt, t0, t1, spos, sgam, s0, s1, tb, torg, tpos, tgam, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < t1) uj_mixed_store(d, perm[s], ldo, U, J, reset);
}

}  // namespace vpmx
