// viscous.hip -- viscous diffusion nu lap(omega) of the particle field
// (docs/src/theory/rvpm.md:182-193,367-368; vpm_viscous,
// src/FLOWUnsteady_simulation.jl:42,229-234), FP64:
//
//   pse_kernel         particle strength exchange (Degond & Mas-Gallic):
//                      V_i = (2 nu / s^2) sum_j (vol_i G_j - vol_j G_i) eta_s(x_i - x_j),
//                      eta(rho) = -zeta'(rho)/rho, s^2 = (sigma_i^2 + sigma_j^2)/2
//                      (symmetric: the sum conserves sum_i G_i).
//   zeta_sum_kernel    vorticity of the regularised particles at the particles,
//                      w_i = sum_j G_j zeta(r_ij / s_j) / s_j^3 (s_j = sigma_j
//                      or a uniform s0), the matrix-vector product of core
//                      spreading's radial-basis-function re-fit.
//   cg_* kernels       the vector steps of the conjugate-gradient solve (three
//                      independent systems, one per component), with fixed-order
//                      block reductions so the iteration is bitwise repeatable.
// Sources stream through LDS in tiles of VB records, one target per lane.
#include "device_math.h"
#include "internal.h"

namespace vpmx {

namespace {

constexpr int VB = 256;
constexpr int VRED = 512;    // partial-sum blocks of the reductions

inline unsigned vgrid(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Source records for PSE: (x, y, z, sigma^2) (Gx, Gy, Gz, vol).
__global__ __launch_bounds__(VB) void pack_pse_kernel(const double *__restrict__ X,
                                                      const double *__restrict__ G,
                                                      const double *__restrict__ sg,
                                                      const double *__restrict__ vol, int64_t n,
                                                      int64_t cap, double4 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n) return;
    const double s = sg[i];
    out[2 * i] = make_double4(X[i], X[cap + i], X[2 * cap + i], s * s);
    out[2 * i + 1] = make_double4(G[i], G[cap + i], G[2 * cap + i], vol[i]);
}

template <int K>
__global__ __launch_bounds__(VB) void pse_kernel(const double4 *__restrict__ src, int64_t n,
                                                 double nu, int64_t cap, double *__restrict__ V)
{
    __shared__ double4 tile[2 * VB];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * VB + t;
    const bool active = i < n;
    double4 a = make_double4(0, 0, 0, 1), b = make_double4(0, 0, 0, 0);
    if (active) {
        a = src[2 * i];
        b = src[2 * i + 1];
    }
    double v0 = 0, v1 = 0, v2 = 0;
    for (int64_t base = 0; base < n; base += VB) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < n) {
            tile[2 * t] = src[2 * js];
            tile[2 * t + 1] = src[2 * js + 1];
        }
        __syncthreads();
        const int cnt = (int)((n - base) < VB ? (n - base) : VB);
        for (int jj = 0; jj < cnt; ++jj) {
            if (base + jj == i) continue;
            const double4 p = tile[2 * jj], q = tile[2 * jj + 1];
            const double d0 = a.x - p.x, d1 = a.y - p.y, d2 = a.z - p.z;
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double s2 = 0.5 * (a.w + p.w);
            const double s = sqrt(s2);
            const double w = 2.0 * nu / s2 * eta_f64<K>(r / s) / (s2 * s);
            v0 += w * (b.w * q.x - q.w * b.x);
            v1 += w * (b.w * q.y - q.w * b.y);
            v2 += w * (b.w * q.z - q.w * b.z);
        }
    }
    if (active) {
        V[i] = v0;
        V[cap + i] = v1;
        V[2 * cap + i] = v2;
    }
}

// Source records for the vorticity sums: (x, y, z, 1/s) (Gx, Gy, Gz, 1/s^3),
// with the strength of masked (static) sources set to zero.
__global__ __launch_bounds__(VB) void pack_zeta_kernel(const double *__restrict__ X,
                                                       const double *__restrict__ G, int64_t gld,
                                                       const double *__restrict__ sg, double s0,
                                                       const uint8_t *__restrict__ st, int64_t n,
                                                       int64_t cap, double4 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n) return;
    const double s = s0 > 0.0 ? s0 : sg[i];
    const double is = 1.0 / s;
    const bool off = st && st[i];
    out[2 * i] = make_double4(X[i], X[cap + i], X[2 * cap + i], is);
    out[2 * i + 1] = make_double4(off ? 0.0 : G[i], off ? 0.0 : G[gld + i],
                                  off ? 0.0 : G[2 * gld + i], is * is * is);
}

template <int K>
__global__ __launch_bounds__(VB) void zeta_sum_kernel(const double4 *__restrict__ src, int64_t n,
                                                      const double *__restrict__ X, int64_t cap,
                                                      double *__restrict__ w, int64_t wld)
{
    __shared__ double4 tile[2 * VB];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * VB + t;
    const bool active = i < n;
    double xi = 0, yi = 0, zi = 0;
    if (active) {
        xi = X[i];
        yi = X[cap + i];
        zi = X[2 * cap + i];
    }
    double w0 = 0, w1 = 0, w2 = 0;
    for (int64_t base = 0; base < n; base += VB) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < n) {
            tile[2 * t] = src[2 * js];
            tile[2 * t + 1] = src[2 * js + 1];
        }
        __syncthreads();
        const int cnt = (int)((n - base) < VB ? (n - base) : VB);
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 p = tile[2 * jj], q = tile[2 * jj + 1];
            const double d0 = xi - p.x, d1 = yi - p.y, d2 = zi - p.z;
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double z = zeta_f64<K>(r * p.w) * q.w;
            w0 += z * q.x;
            w1 += z * q.y;
            w2 += z * q.z;
        }
    }
    if (active) {
        w[i] = w0;
        w[wld + i] = w1;
        w[2 * wld + i] = w2;
    }
}

// Fixed-order block reduction of three sums of products a_k . b_k (masked
// entries excluded); partials of block b at part[3 b + k].
__global__ __launch_bounds__(VB) void cg_dot_partial_kernel(const double *__restrict__ a,
                                                            const double *__restrict__ b,
                                                            const uint8_t *__restrict__ st,
                                                            int64_t n, int64_t ld,
                                                            double *__restrict__ part)
{
    __shared__ double red[3][VB];
    double s0 = 0, s1 = 0, s2 = 0;
    for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
        if (st && st[i]) continue;
        s0 += a[i] * b[i];
        s1 += a[ld + i] * b[ld + i];
        s2 += a[2 * ld + i] * b[2 * ld + i];
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int h = VB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void cg_dot_final_kernel(const double *__restrict__ part, int nb,
                                    double *__restrict__ out)
{
    if (threadIdx.x >= 3) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[3 * b + threadIdx.x];
    out[threadIdx.x] = s;
}

// r = w - Ap, p = r (static entries: 0)
__global__ __launch_bounds__(VB) void cg_init_kernel(const double *__restrict__ w,
                                                     const double *__restrict__ Ap,
                                                     const uint8_t *__restrict__ st, int64_t n,
                                                     int64_t ld, double *__restrict__ r,
                                                     double *__restrict__ p)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n) return;
    const bool off = st && st[i];
    for (int k = 0; k < 3; ++k) {
        const double v = off ? 0.0 : w[k * ld + i] - Ap[k * ld + i];
        r[k * ld + i] = v;
        p[k * ld + i] = v;
    }
}

// G += a_k p (G stride gld), r -= a_k Ap, for the free entries
__global__ __launch_bounds__(VB) void cg_step_kernel(const double *__restrict__ coef,
                                                     const double *__restrict__ p,
                                                     const double *__restrict__ Ap,
                                                     const uint8_t *__restrict__ st, int64_t n,
                                                     int64_t ld, double *__restrict__ G,
                                                     int64_t gld, double *__restrict__ r)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n || (st && st[i])) return;
    for (int k = 0; k < 3; ++k) {
        G[k * gld + i] += coef[k] * p[k * ld + i];
        r[k * ld + i] -= coef[k] * Ap[k * ld + i];
    }
}

// p = r + b_k p for the components still iterating (coef[3 + k] = 1)
__global__ __launch_bounds__(VB) void cg_dir_kernel(const double *__restrict__ coef,
                                                    const double *__restrict__ r,
                                                    const uint8_t *__restrict__ st, int64_t n,
                                                    int64_t ld, double *__restrict__ p)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n || (st && st[i])) return;
    for (int k = 0; k < 3; ++k)
        if (coef[3 + k] != 0.0) p[k * ld + i] = r[k * ld + i] + coef[k] * p[k * ld + i];
}

// Largest free core; partials then one block.
__global__ __launch_bounds__(VB) void max_sigma_partial_kernel(const double *__restrict__ sg,
                                                               const uint8_t *__restrict__ st,
                                                               int64_t n, double *__restrict__ part)
{
    __shared__ double red[VB];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB)
        if (!(st && st[i])) m = fmax(m, sg[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = VB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void max_final_kernel(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double m = 0.0;
    for (int b = 0; b < nb; ++b) m = fmax(m, part[b]);
    out[0] = m;
}

__global__ __launch_bounds__(VB) void set_sigma_kernel(double *__restrict__ sg,
                                                       const uint8_t *__restrict__ st, int64_t n,
                                                       double s0)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n || (st && st[i])) return;
    sg[i] = s0;
}

}  // namespace

hipError_t launch_pse(hipStream_t st, int kernel, const DeviceField &f, double nu, double4 *rec)
{
    if (f.np == 0) return hipSuccess;
    pack_pse_kernel<<<vgrid(f.np, VB), VB, 0, st>>>(f.X, f.G, f.sigma, f.vol, f.np, f.cap, rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned g = vgrid(f.np, VB);
    switch (kernel) {
    case GAUSSIAN: pse_kernel<GAUSSIAN><<<g, VB, 0, st>>>(rec, f.np, nu, f.cap, f.V); break;
    case GAUSSIANERF: pse_kernel<GAUSSIANERF><<<g, VB, 0, st>>>(rec, f.np, nu, f.cap, f.V); break;
    case WINCKELMANS: pse_kernel<WINCKELMANS><<<g, VB, 0, st>>>(rec, f.np, nu, f.cap, f.V); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zeta_sum(hipStream_t st, int kernel, const DeviceField &f, const double *G,
                           int64_t gld, double s0, bool mask_static, double4 *rec, double *w,
                           int64_t wld)
{
    if (f.np == 0) return hipSuccess;
    const uint8_t *stf = (mask_static && f.any_static) ? f.is_static : nullptr;
    pack_zeta_kernel<<<vgrid(f.np, VB), VB, 0, st>>>(f.X, G, gld, f.sigma, s0, stf, f.np, f.cap,
                                                     rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned g = vgrid(f.np, VB);
    switch (kernel) {
    case GAUSSIAN: zeta_sum_kernel<GAUSSIAN><<<g, VB, 0, st>>>(rec, f.np, f.X, f.cap, w, wld); break;
    case GAUSSIANERF: zeta_sum_kernel<GAUSSIANERF><<<g, VB, 0, st>>>(rec, f.np, f.X, f.cap, w, wld); break;
    case WINCKELMANS: zeta_sum_kernel<WINCKELMANS><<<g, VB, 0, st>>>(rec, f.np, f.X, f.cap, w, wld); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cg_dot(hipStream_t st, const double *a, const double *b, const uint8_t *mask,
                         int64_t n, int64_t ld, double *part, double *out3)
{
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(VRED, vgrid(n, VB)));
    cg_dot_partial_kernel<<<nb, VB, 0, st>>>(a, b, mask, n, ld, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    cg_dot_final_kernel<<<1, 64, 0, st>>>(part, nb, out3);
    return hipGetLastError();
}

hipError_t launch_cg_init(hipStream_t st, const double *w, const double *Ap, const uint8_t *mask,
                          int64_t n, int64_t ld, double *r, double *p)
{
    if (n == 0) return hipSuccess;
    cg_init_kernel<<<vgrid(n, VB), VB, 0, st>>>(w, Ap, mask, n, ld, r, p);
    return hipGetLastError();
}

hipError_t launch_cg_step(hipStream_t st, const double *coef, const double *p, const double *Ap,
                          const uint8_t *mask, int64_t n, int64_t ld, double *G, int64_t gld,
                          double *r)
{
    if (n == 0) return hipSuccess;
    cg_step_kernel<<<vgrid(n, VB), VB, 0, st>>>(coef, p, Ap, mask, n, ld, G, gld, r);
    return hipGetLastError();
}

hipError_t launch_cg_dir(hipStream_t st, const double *coef, const double *r, const uint8_t *mask,
                         int64_t n, int64_t ld, double *p)
{
    if (n == 0) return hipSuccess;
    cg_dir_kernel<<<vgrid(n, VB), VB, 0, st>>>(coef, r, mask, n, ld, p);
    return hipGetLastError();
}

hipError_t launch_max_sigma(hipStream_t st, const DeviceField &f, double *part, double *out)
{
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(VRED, vgrid(f.np, VB)));
    max_sigma_partial_kernel<<<nb, VB, 0, st>>>(f.sigma, f.any_static ? f.is_static : nullptr,
                                                f.np, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    max_final_kernel<<<1, 64, 0, st>>>(part, nb, out);
    return hipGetLastError();
}

hipError_t launch_set_sigma(hipStream_t st, const DeviceField &f, double s0)
{
    if (f.np == 0) return hipSuccess;
    set_sigma_kernel<<<vgrid(f.np, VB), VB, 0, st>>>(f.sigma, f.any_static ? f.is_static : nullptr,
                                                     f.np, s0);
    return hipGetLastError();
}

}  // namespace vpmx
