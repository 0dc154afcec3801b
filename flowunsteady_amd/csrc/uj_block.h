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
// extended to J = grad U.  `tile` holds 2*BLK records.
template <int K, int BLK>
__device__ __forceinline__ void uj_f64_block(const double4 *__restrict__ src, int64_t s0,
                                             int64_t s1, const double *__restrict__ X,
                                             int64_t t0, int64_t t1, int64_t cap,
                                             double *__restrict__ U, double *__restrict__ J,
                                             int reset, double4 *tile)
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
    if (active) {
        if (!reset) {
            u0 += U[i]; u1 += U[cap + i]; u2 += U[2 * cap + i];
            j0 += J[i]; j1 += J[cap + i]; j2 += J[2 * cap + i];
            j3 += J[3 * cap + i]; j4 += J[4 * cap + i]; j5 += J[5 * cap + i];
            j6 += J[6 * cap + i]; j7 += J[7 * cap + i]; j8 += J[8 * cap + i];
        }
        U[i] = u0; U[cap + i] = u1; U[2 * cap + i] = u2;
        J[i] = j0; J[cap + i] = j1; J[2 * cap + i] = j2;
        J[3 * cap + i] = j3; J[4 * cap + i] = j4; J[5 * cap + i] = j5;
        J[6 * cap + i] = j6; J[7 * cap + i] = j7; J[8 * cap + i] = j8;
    }
}

// Mixed precision.  Targets are the sorted slots [t0, t1) of `tgt`; sources
// the sorted slots [s0, s1) of (spos, sgam).  Coordinates are taken relative
// to the block's first target in FP64 and rounded to FP32; each tile's FP32
// partials are folded into FP64.  Results go to the target's own index
// perm[slot] (stride ldo).  `tpos`, `tgam` hold BLK records each.
template <int K, int BLK>
__device__ __forceinline__ void uj_mixed_block(const double4 *__restrict__ tgt, int64_t t0,
                                               int64_t t1, const int32_t *__restrict__ perm,
                                               const double4 *__restrict__ spos,
                                               const float4 *__restrict__ sgam, int64_t s0,
                                               int64_t s1, int64_t ldo,
                                               double *__restrict__ U, double *__restrict__ J,
                                               int reset, float4 *tpos, float4 *tgam)
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
    double dU0 = 0, dU1 = 0, dU2 = 0, dW0 = 0, dW1 = 0, dW2 = 0;
    double dJ0 = 0, dJ1 = 0, dJ2 = 0, dJ3 = 0, dJ4 = 0, dJ5 = 0, dJ6 = 0, dJ7 = 0;

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
        MixedAcc acc;
        if (cnt == BLK) {
#pragma unroll 4
            for (int jj = 0; jj < BLK; ++jj)
                mixed_pair<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
        } else {
            for (int jj = 0; jj < cnt; ++jj)
                mixed_pair<K>(tpos[jj], tgam[jj], xi, yi, zi, acc);
        }
        dU0 += acc.U0; dU1 += acc.U1; dU2 += acc.U2;
        dW0 += acc.W0; dW1 += acc.W1; dW2 += acc.W2;
        dJ0 += acc.J0; dJ1 += acc.J1; dJ2 += acc.J2; dJ3 += acc.J3;
        dJ4 += acc.J4; dJ5 += acc.J5; dJ6 += acc.J6; dJ7 += acc.J7;
    }
    if (active) {
        const int64_t i = perm[s];
        const double k = -INV4PI;
        double u0 = k * dU0, u1 = k * dU1, u2 = k * dU2;
        // J = -1/(4 pi) (radial + eps_klm W_m); J22 from the zero trace
        double j0 = k * dJ0, j1 = k * (dJ1 - dW2), j2 = k * (dJ2 + dW1);
        double j3 = k * (dJ3 + dW2), j4 = k * dJ4, j5 = k * (dJ5 - dW0);
        double j6 = k * (dJ6 - dW1), j7 = k * (dJ7 + dW0), j8 = -k * (dJ0 + dJ4);
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
}

}  // namespace vpmx
