// kernels.hip -- CDNA4 (gfx950) kernels of the rVPM particle-interaction step.
//
//   uj_f64_kernel      all-pairs U and J in FP64, reference pair algebra
//                      (src/FLOWUnsteady_processing_force.jl:888-905 extended to
//                      J = grad U); LDS-tiled sources, one target per lane.
//   uj_mixed_kernel    all-pairs U and J with FP32 pair math on Morton-sorted
//                      targets, coordinates relative to the block's first
//                      target as FP32 (hi, lo) pairs (lo used within 7
//                      cores), FP32 tile partials folded into FP64 per tile.
//   estr_f64_kernel    SFS stretching Estr (docs/src/theory/rvpm.md:251-262).
//   update_kernel      fused low-storage RK / Euler substep of the rVPM
//                      equations (rvpm.md:107-178,196-236,364) in FP64.
//   relax_kernel       Pedrizzetti relaxation (rvpm.md:363).
//   diag kernels       enstrophy 0.5 sum Gamma.omega (src/FLOWUnsteady_monitors.jl:573)
//                      and energy sum U.(x x Gamma).
#include <rocprim/device/device_radix_sort.hpp>

#include "device_math.h"
#include "internal.h"
#include "uj_block.h"

namespace vpmx {

constexpr int UJ_BLOCK = 256;
constexpr int EW_BLOCK = 256;      // elementwise kernels
constexpr int RED_BLOCKS = 1024;   // partial-sum blocks of the reductions

static inline unsigned grid_for(int64_t n, int block)
{
    return (unsigned)((n + block - 1) / block);
}

// ===========================================================================
// FP64 path
// ===========================================================================

// Source record (x, y, z, Gx) (Gy, Gz, sigma, 1/sigma).
__global__ __launch_bounds__(EW_BLOCK) void pack_f64_kernel(const double *__restrict__ X,
                                                            const double *__restrict__ G,
                                                            const double *__restrict__ sg,
                                                            int64_t n, int64_t cap,
                                                            double4 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    const double s = sg[i];
    out[2 * i] = make_double4(X[i], X[cap + i], X[2 * cap + i], G[i]);
    out[2 * i + 1] = make_double4(G[cap + i], G[2 * cap + i], s, 1.0 / s);
}

template <int K, bool DS>
__global__ __launch_bounds__(UJ_BLOCK) void uj_f64_kernel(const double4 *__restrict__ src,
                                                          int64_t ns,
                                                          const double *__restrict__ X,
                                                          int64_t nt, int64_t cap,
                                                          double *__restrict__ U,
                                                          double *__restrict__ J, int reset)
{
    __shared__ double4 tile[2 * UJ_BLOCK];
    const int64_t t0 = (int64_t)blockIdx.x * UJ_BLOCK;
    uj_f64_block<K, UJ_BLOCK, DS>(src, 0, ns, X, t0, nt, cap, U, J, reset, tile);
}

// Split source range: workgroup (block b, segment g) sums sources
// [g seglen, (g+1) seglen) for target block b into part[(g 12 + k) pld + i].
template <int K, bool DS>
__global__ __launch_bounds__(UJ_BLOCK) void uj_f64_split_kernel(const double4 *__restrict__ src,
                                                                int64_t ns,
                                                                const double *__restrict__ X,
                                                                int64_t nt, int64_t cap,
                                                                int64_t seglen, int nseg,
                                                                double *__restrict__ part,
                                                                int64_t pld)
{
    __shared__ double4 tile[2 * UJ_BLOCK];
    const int g = (int)(blockIdx.x % nseg);
    const int64_t t0 = (int64_t)(blockIdx.x / nseg) * UJ_BLOCK;
    const int64_t s0 = (int64_t)g * seglen;
    const int64_t s1 = s0 + seglen < ns ? s0 + seglen : ns;
    double d[12];
    uj_f64_sums<K, UJ_BLOCK, DS>(src, s0, s1, X, t0, nt, cap, tile, d);
    const int64_t i = t0 + threadIdx.x;
    if (i < nt)
        for (int k = 0; k < 12; ++k) part[((int64_t)g * 12 + k) * pld + i] = d[k];
}

// Segments of each target added in segment order (bitwise repeatable).
__global__ __launch_bounds__(EW_BLOCK) void uj_f64_split_reduce_kernel(const double *__restrict__ part,
                                                                       int64_t pld, int nseg,
                                                                       int64_t nt, int64_t ldo,
                                                                       double *__restrict__ U,
                                                                       double *__restrict__ J,
                                                                       int reset)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= nt) return;
    double d[12];
    for (int k = 0; k < 12; ++k) {
        double v = 0.0;
        for (int g = 0; g < nseg; ++g) v += part[((int64_t)g * 12 + k) * pld + i];
        d[k] = v;
    }
    uj_f64_store(d, i, ldo, U, J, reset);
}

// How a launch over nb target blocks and n sources is split: when the target
// blocks alone cannot fill the device, the source range is cut into segments
// (multiples of UJ_BLOCK, at least four tiles each) so that about 1024
// workgroups run; their partial sums go through `part` (SPLIT_PART doubles).
struct Split {
    int nseg;
    int64_t seglen;
};

static Split split_for(int64_t nb, int64_t n)
{
    if (nb >= 512 || n <= 4 * UJ_BLOCK) return {1, n};
    int64_t nseg = std::min<int64_t>((1024 + nb - 1) / nb, (n + 4 * UJ_BLOCK - 1) / (4 * UJ_BLOCK));
    nseg = std::max<int64_t>(nseg, 1);
    const int64_t seglen = ((n + nseg - 1) / nseg + UJ_BLOCK - 1) / UJ_BLOCK * UJ_BLOCK;
    return {(int)((n + seglen - 1) / seglen), seglen};
}

hipError_t launch_pack_f64(hipStream_t st, const DeviceField &f, Scratch &s)
{
    if (f.np == 0) return hipSuccess;
    pack_f64_kernel<<<grid_for(f.np, EW_BLOCK), EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, f.np, f.cap,
                                                                   s.src64);
    return hipGetLastError();
}

template <int K, bool DS = false>
static hipError_t uj_f64_go(hipStream_t st, const double4 *src, int64_t ns, const double *X,
                            int64_t nt, int64_t cap, double *U, double *J, int reset,
                            double *part)
{
    const int64_t nb = (nt + UJ_BLOCK - 1) / UJ_BLOCK;
    const Split sp = split_for(nb, ns);
    if (sp.nseg == 1 || !part) {
        uj_f64_kernel<K, DS><<<(unsigned)nb, UJ_BLOCK, 0, st>>>(src, ns, X, nt, cap, U, J, reset);
        return hipGetLastError();
    }
    uj_f64_split_kernel<K, DS><<<(unsigned)(nb * sp.nseg), UJ_BLOCK, 0, st>>>(src, ns, X, nt, cap,
                                                                          sp.seglen, sp.nseg,
                                                                          part, nt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uj_f64_split_reduce_kernel<<<grid_for(nt, EW_BLOCK), EW_BLOCK, 0, st>>>(part, nt, sp.nseg, nt,
                                                                           cap, U, J, reset);
    return hipGetLastError();
}

static hipError_t uj_f64_any(hipStream_t st, int kernel, const double4 *src, int64_t ns,
                             const double *X, int64_t nt, int64_t cap, double *U, double *J,
                             int reset, double *part)
{
    switch (kernel) {
    case SINGULAR: return uj_f64_go<SINGULAR>(st, src, ns, X, nt, cap, U, J, reset, part);
    case GAUSSIAN: return uj_f64_go<GAUSSIAN>(st, src, ns, X, nt, cap, U, J, reset, part);
    case GAUSSIANERF: return uj_f64_go<GAUSSIANERF>(st, src, ns, X, nt, cap, U, J, reset, part);
    case WINCKELMANS: return uj_f64_go<WINCKELMANS>(st, src, ns, X, nt, cap, U, J, reset, part);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_uj_f64_targets(hipStream_t st, int kernel, const DeviceField &f,
                                 const Scratch &s, const double *tX, int64_t nt, double *U,
                                 double *J)
{
    if (nt == 0 || f.np == 0) return hipSuccess;
    return uj_f64_any(st, kernel, s.src64, f.np, tX, nt, nt, U, J, 1, s.kpart);
}

hipError_t launch_uj_f64(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                         int reset)
{
    if (f.np == 0) return hipSuccess;
    return uj_f64_any(st, kernel, s.src64, f.np, f.X, f.np, f.cap, f.U, f.J, reset, s.kpart);
}

// ---- FP64 on the Morton-sorted field (single fields) -----------------------

// Sorted source records: slot s holds particle perm[s] as
// (x, y, z, Gx) (Gy, Gz, sigma, 1/sigma).
__global__ __launch_bounds__(EW_BLOCK) void pack_f64_sorted_kernel(const double *__restrict__ X,
                                                                   const double *__restrict__ G,
                                                                   const double *__restrict__ sg,
                                                                   const int32_t *__restrict__ perm,
                                                                   int64_t n, int64_t cap,
                                                                   double4 *__restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm[s];
    const double sig = sg[i];
    out[2 * s] = make_double4(X[i], X[cap + i], X[2 * cap + i], G[i]);
    out[2 * s + 1] = make_double4(G[cap + i], G[2 * cap + i], sig, 1.0 / sig);
}

// Bounds of each 256-slot tile of the sorted records: bbox centre, half
// diagonal, largest sigma.  One workgroup per tile.
__global__ __launch_bounds__(UJ_BLOCK) void tile_bounds_f64_kernel(const double4 *__restrict__ src,
                                                                   int64_t n,
                                                                   double4 *__restrict__ tb)
{
    __shared__ double red[7][UJ_BLOCK];
    const int t = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.x * UJ_BLOCK + t;
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    if (s < n) {
        const double4 p = src[2 * s];
        v[0] = v[3] = p.x;
        v[1] = v[4] = p.y;
        v[2] = v[5] = p.z;
        v[6] = src[2 * s + 1].z;
    }
    for (int k = 0; k < 7; ++k) red[k][t] = v[k];
    __syncthreads();
    for (int w = UJ_BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 3; ++k) red[k][t] = fmin(red[k][t], red[k][t + w]);
            for (int k = 3; k < 7; ++k) red[k][t] = fmax(red[k][t], red[k][t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const double hx = 0.5 * (red[3][0] - red[0][0]), hy = 0.5 * (red[4][0] - red[1][0]),
                     hz = 0.5 * (red[5][0] - red[2][0]);
        tb[2 * blockIdx.x] = make_double4(red[0][0] + hx, red[1][0] + hy, red[2][0] + hz,
                                          sqrt(hx * hx + hy * hy + hz * hz));
        tb[2 * blockIdx.x + 1] = make_double4(red[6][0], 0.0, 0.0, 0.0);
    }
}

template <int K>
__global__ __launch_bounds__(UJ_BLOCK) void uj_f64s_kernel(const double4 *__restrict__ src,
                                                           int64_t n,
                                                           const double4 *__restrict__ tb,
                                                           const int32_t *__restrict__ perm,
                                                           int64_t cap, double *__restrict__ U,
                                                           double *__restrict__ J, int reset)
{
    __shared__ double4 tile[2 * UJ_BLOCK];
    const int64_t t0 = (int64_t)blockIdx.x * UJ_BLOCK;
    double d[15];
    uj_f64s_sums<K, UJ_BLOCK>(src, 0, n, t0, n, tb, tile, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < n) uj_f64s_store(d, perm[s], cap, U, J, reset);
}

// Split source range: workgroup (block b, segment g) writes its 15 sums to
// part[(g 15 + k) pld + slot].
template <int K>
__global__ __launch_bounds__(UJ_BLOCK) void uj_f64s_split_kernel(const double4 *__restrict__ src,
                                                                 int64_t n,
                                                                 const double4 *__restrict__ tb,
                                                                 int64_t seglen, int nseg,
                                                                 double *__restrict__ part,
                                                                 int64_t pld)
{
    __shared__ double4 tile[2 * UJ_BLOCK];
    const int g = (int)(blockIdx.x % nseg);
    const int64_t t0 = (int64_t)(blockIdx.x / nseg) * UJ_BLOCK;
    const int64_t s0 = (int64_t)g * seglen;
    const int64_t s1 = s0 + seglen < n ? s0 + seglen : n;
    double d[15];
    uj_f64s_sums<K, UJ_BLOCK>(src, s0, s1, t0, n, tb, tile, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < n)
        for (int k = 0; k < 15; ++k) part[((int64_t)g * 15 + k) * pld + s] = d[k];
}

// Segments of each slot added in segment order (bitwise repeatable), stored
// through perm.
__global__ __launch_bounds__(EW_BLOCK) void uj_f64s_split_reduce_kernel(const double *__restrict__ part,
                                                                        int64_t pld, int nseg,
                                                                        int64_t n,
                                                                        const int32_t *__restrict__ perm,
                                                                        int64_t ldo,
                                                                        double *__restrict__ U,
                                                                        double *__restrict__ J,
                                                                        int reset)
{
    const int64_t s = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (s >= n) return;
    double d[15];
    for (int k = 0; k < 15; ++k) {
        double v = 0.0;
        for (int g = 0; g < nseg; ++g) v += part[((int64_t)g * 15 + k) * pld + s];
        d[k] = v;
    }
    uj_f64s_store(d, perm[s], ldo, U, J, reset);
}

hipError_t launch_pack_f64_sorted(hipStream_t st, const DeviceField &f, Scratch &s)
{
    const int64_t n = f.np;
    if (n == 0) return hipSuccess;
    pack_f64_sorted_kernel<<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, s.perm, n,
                                                                       f.cap, s.src64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    tile_bounds_f64_kernel<<<grid_for(n, UJ_BLOCK), UJ_BLOCK, 0, st>>>(s.src64, n, s.tbnd);
    return hipGetLastError();
}

template <int K>
static hipError_t uj_f64s_go(hipStream_t st, const DeviceField &f, const Scratch &s, int reset)
{
    const int64_t n = f.np;
    const int64_t nb = (n + UJ_BLOCK - 1) / UJ_BLOCK;
    const Split sp = split_for(nb, n);
    if (sp.nseg == 1 || !s.kpart || (int64_t)sp.nseg * 15 * n > s.kcap) {
        uj_f64s_kernel<K><<<(unsigned)nb, UJ_BLOCK, 0, st>>>(s.src64, n, s.tbnd, s.perm, f.cap,
                                                             f.U, f.J, reset);
        return hipGetLastError();
    }
    uj_f64s_split_kernel<K><<<(unsigned)(nb * sp.nseg), UJ_BLOCK, 0, st>>>(
        s.src64, n, s.tbnd, sp.seglen, sp.nseg, s.kpart, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uj_f64s_split_reduce_kernel<<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(
        s.kpart, n, sp.nseg, n, s.perm, f.cap, f.U, f.J, reset);
    return hipGetLastError();
}

hipError_t launch_uj_f64_sorted(hipStream_t st, int kernel, const DeviceField &f,
                                const Scratch &s, int reset)
{
    if (f.np == 0) return hipSuccess;
    switch (kernel) {
    case SINGULAR: return uj_f64s_go<SINGULAR>(st, f, s, reset);
    case GAUSSIAN: return uj_f64s_go<GAUSSIAN>(st, f, s, reset);
    case GAUSSIANERF: return uj_f64s_go<GAUSSIANERF>(st, f, s, reset);
    case WINCKELMANS:
        // no far-field form (1 - g decays as rho^-4): the unsorted-target
        // kernel over the sorted sources keeps its higher occupancy
        return uj_f64_go<WINCKELMANS>(st, s.src64, f.np, f.X, f.np, f.cap, f.U, f.J, reset,
                                      s.kpart);
    default: return hipErrorInvalidValue;
    }
}

// Caller-owned device arrays (stride n) and a caller-provided scratch block:
// the source records (2n double4) first, then the split partials when the
// launch is split (12 doubles per target per segment).
static size_t uj_f64_dev_part_offset(int64_t n)
{
    return ((size_t)n * 2 * sizeof(double4) + 255) / 256 * 256;
}

size_t uj_f64_dev_scratch_bytes(int64_t n)
{
    const Split sp = split_for((n + UJ_BLOCK - 1) / UJ_BLOCK, n);
    const size_t part = sp.nseg > 1 ? (size_t)sp.nseg * 12 * (size_t)n * sizeof(double) : 0;
    return uj_f64_dev_part_offset(n) + part;
}

hipError_t launch_uj_f64_dev(hipStream_t st, int kernel, int64_t n, const double *X,
                             const double *G, const double *sigma, double *U, double *J,
                             int reset, void *scratch)
{
    if (n == 0) return hipSuccess;
    double4 *src = static_cast<double4 *>(scratch);
    double *part = reinterpret_cast<double *>(static_cast<char *>(scratch) +
                                              uj_f64_dev_part_offset(n));
    pack_f64_kernel<<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(X, G, sigma, n, n, src);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return uj_f64_any(st, kernel, src, n, X, n, n, U, J, reset, part);
}

// Core-size derivatives D[U], D[J] of every particle into DU (3 x cap) and DJ
// (9 x cap), FP64, from the source records of the UJ that just ran.
hipError_t launch_dsig_f64(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                           double *DU, double *DJ)
{
    if (f.np == 0) return hipSuccess;
    switch (kernel) {
    case GAUSSIAN: return uj_f64_go<GAUSSIAN, true>(st, s.src64, f.np, f.X, f.np, f.cap, DU, DJ, 1, s.kpart);
    case GAUSSIANERF: return uj_f64_go<GAUSSIANERF, true>(st, s.src64, f.np, f.X, f.np, f.cap, DU, DJ, 1, s.kpart);
    case WINCKELMANS: return uj_f64_go<WINCKELMANS, true>(st, s.src64, f.np, f.X, f.np, f.cap, DU, DJ, 1, s.kpart);
    default: return hipErrorInvalidValue;
    }
}

// ===========================================================================
// Mixed path: Morton sort, packing, FP32 pair kernel
// ===========================================================================

__global__ __launch_bounds__(EW_BLOCK) void bbox_partial_kernel(const double *__restrict__ X,
                                                                int64_t n, int64_t cap,
                                                                double *__restrict__ part)
{
    __shared__ double red[6][EW_BLOCK];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * EW_BLOCK) {
        for (int k = 0; k < 3; ++k) {
            const double v = X[k * cap + i];
            lo[k] = fmin(lo[k], v);
            hi[k] = fmax(hi[k], v);
        }
    }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = lo[k];
        red[3 + k][threadIdx.x] = hi[k];
    }
    __syncthreads();
    for (int w = EW_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Reduces the bbox partials (every block redoes the same fixed-order reduction
// of `nparts` partials, which is cheap) and writes 30-bit Morton keys.
__global__ __launch_bounds__(EW_BLOCK) void morton_kernel(const double *__restrict__ X,
                                                          int64_t n, int64_t cap,
                                                          const double *__restrict__ part,
                                                          int nparts, uint32_t *__restrict__ key,
                                                          int32_t *__restrict__ val)
{
    __shared__ double box[6];
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double v = part[k];
        for (int b = 1; b < nparts; ++b)
            v = (k < 3) ? fmin(v, part[6 * b + k]) : fmax(v, part[6 * b + k]);
        box[k] = v;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    key[i] = morton30(box, X[i], X[cap + i], X[2 * cap + i]);
    val[i] = (int32_t)i;
}

// Per-source constants of the FP32 radial functions: gaussian(-erf)
// p0 = 1/sigma; winckelmans p0 = sigma^2; p1 = 1/sigma^2 for every kernel.
template <int K>
__global__ __launch_bounds__(EW_BLOCK) void pack_mixed_kernel(const double *__restrict__ X,
                                                              const double *__restrict__ G,
                                                              const double *__restrict__ sg,
                                                              const int32_t *__restrict__ perm,
                                                              int64_t n, int64_t cap,
                                                              double4 *__restrict__ spos,
                                                              float4 *__restrict__ sgam)
{
    const int64_t s = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm[s];
    const double sig = sg[i];
    double p0 = 0.0;
    const double p1 = 1.0 / (sig * sig);
    if constexpr (K == GAUSSIANERF || K == GAUSSIAN) {
        p0 = 1.0 / sig;
    } else if constexpr (K == WINCKELMANS) {
        p0 = sig * sig;
    }
    spos[s] = make_double4(X[i], X[cap + i], X[2 * cap + i], p0);
    sgam[s] = make_float4((float)G[i], (float)G[cap + i], (float)G[2 * cap + i], (float)p1);
}

template <int K, bool DS>
__global__ __launch_bounds__(UJ_BLOCK) void uj_mixed_kernel(const double4 *__restrict__ tgt,
                                                            int64_t nt,
                                                            const int32_t *__restrict__ perm,
                                                            const double4 *__restrict__ spos,
                                                            const float4 *__restrict__ sgam,
                                                            int64_t n, int64_t cap,
                                                            double *__restrict__ U,
                                                            double *__restrict__ J, int reset,
                                                            const double4 *__restrict__ tb)
{
    __shared__ float4 tpos[UJ_BLOCK];
    __shared__ float4 tgam[UJ_BLOCK];
    __shared__ float4 tlo[UJ_BLOCK];
    const int64_t t0 = (int64_t)blockIdx.x * UJ_BLOCK;
    uj_mixed_block<K, UJ_BLOCK, DS>(tgt, t0, nt, perm, spos, sgam, 0, n, cap, U, J, reset, tpos,
                                    tgam, tlo, tb);
}

// Split source range of the mixed kernel: workgroup (block b, segment g) writes
// its 14 sums to part[(g 14 + k) pld + slot].
template <int K, bool DS>
__global__ __launch_bounds__(UJ_BLOCK) void uj_mixed_split_kernel(const double4 *__restrict__ tgt,
                                                                  int64_t nt,
                                                                  const double4 *__restrict__ spos,
                                                                  const float4 *__restrict__ sgam,
                                                                  int64_t n, int64_t seglen,
                                                                  int nseg,
                                                                  const double4 *__restrict__ tb,
                                                                  double *__restrict__ part,
                                                                  int64_t pld)
{
    __shared__ float4 tpos[UJ_BLOCK];
    __shared__ float4 tgam[UJ_BLOCK];
    __shared__ float4 tlo[UJ_BLOCK];
    const int g = (int)(blockIdx.x % nseg);
    const int64_t t0 = (int64_t)(blockIdx.x / nseg) * UJ_BLOCK;
    const int64_t s0 = (int64_t)g * seglen;
    const int64_t s1 = s0 + seglen < n ? s0 + seglen : n;
    double d[14];
    uj_mixed_sums<K, UJ_BLOCK, DS>(tgt, t0, nt, spos, sgam, s0, s1, tb, 0, tpos, tgam, tlo, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < nt)
        for (int k = 0; k < 14; ++k) part[((int64_t)g * 14 + k) * pld + s] = d[k];
}

__global__ __launch_bounds__(EW_BLOCK) void uj_mixed_split_reduce_kernel(
    const double *__restrict__ part, int64_t pld, int nseg, int64_t nt,
    const int32_t *__restrict__ perm, int64_t ldo, double *__restrict__ U, double *__restrict__ J,
    int reset)
{
    const int64_t s = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (s >= nt) return;
    double d[14];
    for (int k = 0; k < 14; ++k) {
        double v = 0.0;
        for (int g = 0; g < nseg; ++g) v += part[((int64_t)g * 14 + k) * pld + s];
        d[k] = v;
    }
    uj_mixed_store(d, perm[s], ldo, U, J, reset);
}

// Target-subset variant (probes): targets need not be spatially compact, so
// both target and source carry their offset from the block origin as an FP32
// (hi, lo) pair and dX = (t_hi - s_hi) + (t_lo - s_lo); the hi difference is
// exact for the nearby pairs that matter, leaving dX accurate to ~1e-15
// absolute for any block extent.  Sums of target block t0 over [s0, s1).
template <int K>
__device__ __forceinline__ void uj_mixed2_sums(const double4 *__restrict__ tgt, int64_t t0,
                                               int64_t nt, const double4 *__restrict__ spos,
                                               const float4 *__restrict__ sgam, int64_t s0,
                                               int64_t s1, float4 *shi, float4 *slo, float4 *sgm,
                                               double d[14])
{
    const int t = threadIdx.x;
    const int64_t s = t0 + t;
    const bool active = s < nt;
    const double4 o = tgt[t0];
    float xh = 0.f, yh = 0.f, zh = 0.f, xl = 0.f, yl = 0.f, zl = 0.f;
    if (active) {
        const double4 p = tgt[s];
        const double dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
        xh = (float)dx; yh = (float)dy; zh = (float)dz;
        xl = (float)(dx - xh); yl = (float)(dy - yh); zl = (float)(dz - zh);
    }
    for (int k = 0; k < 14; ++k) d[k] = 0.0;
    for (int64_t base = s0; base < s1; base += UJ_BLOCK) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < s1) {
            const double4 p = spos[js];
            const double dx = p.x - o.x, dy = p.y - o.y, dz = p.z - o.z;
            const float hx = (float)dx, hy = (float)dy, hz = (float)dz;
            shi[t] = make_float4(hx, hy, hz, (float)p.w);
            slo[t] = make_float4((float)(dx - hx), (float)(dy - hy), (float)(dz - hz), 0.f);
            sgm[t] = sgam[js];
        }
        __syncthreads();
        const int cnt = (int)((s1 - base) < UJ_BLOCK ? (s1 - base) : UJ_BLOCK);
        MixedAcc acc;
        for (int jj = 0; jj < cnt; ++jj) {
            const float4 h = shi[jj], l = slo[jj];
            mixed_pair_d<K>((xh - h.x) + (xl - l.x), (yh - h.y) + (yl - l.y),
                            (zh - h.z) + (zl - l.z), h.w, sgm[jj], acc);
            if ((jj & (MIXED_FOLD - 1)) == MIXED_FOLD - 1) mixed_fold(d, acc);
        }
        mixed_fold(d, acc);
    }
}

template <int K>
__global__ __launch_bounds__(UJ_BLOCK) void uj_mixed2_kernel(const double4 *__restrict__ tgt,
                                                             int64_t nt,
                                                             const int32_t *__restrict__ perm,
                                                             const double4 *__restrict__ spos,
                                                             const float4 *__restrict__ sgam,
                                                             int64_t n, int64_t ldo,
                                                             double *__restrict__ U,
                                                             double *__restrict__ J)
{
    __shared__ float4 shi[UJ_BLOCK];
    __shared__ float4 slo[UJ_BLOCK];
    __shared__ float4 sgm[UJ_BLOCK];
    const int64_t t0 = (int64_t)blockIdx.x * UJ_BLOCK;
    double d[14];
    uj_mixed2_sums<K>(tgt, t0, nt, spos, sgam, 0, n, shi, slo, sgm, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < nt) uj_mixed_store(d, perm[s], ldo, U, J, 1);
}

template <int K>
__global__ __launch_bounds__(UJ_BLOCK) void uj_mixed2_split_kernel(const double4 *__restrict__ tgt,
                                                                   int64_t nt,
                                                                   const double4 *__restrict__ spos,
                                                                   const float4 *__restrict__ sgam,
                                                                   int64_t n, int64_t seglen,
                                                                   int nseg,
                                                                   double *__restrict__ part,
                                                                   int64_t pld)
{
    __shared__ float4 shi[UJ_BLOCK];
    __shared__ float4 slo[UJ_BLOCK];
    __shared__ float4 sgm[UJ_BLOCK];
    const int g = (int)(blockIdx.x % nseg);
    const int64_t t0 = (int64_t)(blockIdx.x / nseg) * UJ_BLOCK;
    const int64_t s0 = (int64_t)g * seglen;
    const int64_t s1 = s0 + seglen < n ? s0 + seglen : n;
    double d[14];
    uj_mixed2_sums<K>(tgt, t0, nt, spos, sgam, s0, s1, shi, slo, sgm, d);
    const int64_t s = t0 + threadIdx.x;
    if (s < nt)
        for (int k = 0; k < 14; ++k) part[((int64_t)g * 14 + k) * pld + s] = d[k];
}

// Mixed launches.  two = 1: the two-float target kernel (probes).  tb: tile
// bounds of the sorted field when the targets are its own slots (self UJ).
// DS: the core-size derivatives (self UJ only).
template <int K, bool DS = false>
static hipError_t uj_mixed_go(hipStream_t st, bool two, const double4 *tgt, int64_t nt,
                              const int32_t *tperm, const double4 *spos, const float4 *sgam,
                              int64_t n, int64_t ldo, double *U, double *J, int reset,
                              const double4 *tb, double *part)
{
    const int64_t nb = (nt + UJ_BLOCK - 1) / UJ_BLOCK;
    const Split sp = split_for(nb, n);
    if (sp.nseg == 1 || !part) {
        if (two)
            uj_mixed2_kernel<K><<<(unsigned)nb, UJ_BLOCK, 0, st>>>(tgt, nt, tperm, spos, sgam, n,
                                                                  ldo, U, J);
        else
            uj_mixed_kernel<K, DS><<<(unsigned)nb, UJ_BLOCK, 0, st>>>(tgt, nt, tperm, spos, sgam, n,
                                                                     ldo, U, J, reset, tb);
        return hipGetLastError();
    }
    const unsigned g = (unsigned)(nb * sp.nseg);
    if (two)
        uj_mixed2_split_kernel<K><<<g, UJ_BLOCK, 0, st>>>(tgt, nt, spos, sgam, n, sp.seglen,
                                                         sp.nseg, part, nt);
    else
        uj_mixed_split_kernel<K, DS><<<g, UJ_BLOCK, 0, st>>>(tgt, nt, spos, sgam, n, sp.seglen,
                                                            sp.nseg, tb, part, nt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uj_mixed_split_reduce_kernel<<<grid_for(nt, EW_BLOCK), EW_BLOCK, 0, st>>>(
        part, nt, sp.nseg, nt, tperm, ldo, U, J, two ? 1 : reset);
    return hipGetLastError();
}

static hipError_t uj_mixed_any(hipStream_t st, int kernel, bool two, const double4 *tgt,
                               int64_t nt, const int32_t *tperm, const Scratch &s, int64_t n,
                               int64_t ldo, double *U, double *J, int reset, const double4 *tb)
{
    switch (kernel) {
    case SINGULAR:
        return uj_mixed_go<SINGULAR>(st, two, tgt, nt, tperm, s.spos, s.sgam, n, ldo, U, J, reset, tb, s.kpart);
    case GAUSSIAN:
        return uj_mixed_go<GAUSSIAN>(st, two, tgt, nt, tperm, s.spos, s.sgam, n, ldo, U, J, reset, tb, s.kpart);
    case GAUSSIANERF:
        return uj_mixed_go<GAUSSIANERF>(st, two, tgt, nt, tperm, s.spos, s.sgam, n, ldo, U, J, reset, tb, s.kpart);
    case WINCKELMANS:
        return uj_mixed_go<WINCKELMANS>(st, two, tgt, nt, tperm, s.spos, s.sgam, n, ldo, U, J, reset, tb, s.kpart);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_uj_mixed2_targets(hipStream_t st, int kernel, const DeviceField &f,
                                    const Scratch &s, const double4 *tgt, int64_t nt,
                                    const int32_t *tperm, int64_t ldo, double *U, double *J)
{
    if (nt == 0 || f.np == 0) return hipSuccess;
    return uj_mixed_any(st, kernel, true, tgt, nt, tperm, s, f.np, ldo, U, J, 1, nullptr);
}

hipError_t launch_bbox(hipStream_t st, const DeviceField &f, double *part, int *nparts)
{
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, grid_for(f.np, EW_BLOCK)));
    *nparts = nb;
    bbox_partial_kernel<<<nb, EW_BLOCK, 0, st>>>(f.X, f.np, f.cap, part);
    return hipGetLastError();
}

size_t sort_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, bytes, (const uint32_t *)nullptr,
                              (uint32_t *)nullptr, (const int32_t *)nullptr, (int32_t *)nullptr,
                              (unsigned int)n, 0, 30, (hipStream_t)0);
    return bytes;
}

hipError_t launch_sort_mixed(hipStream_t st, const DeviceField &f, Scratch &s)
{
    const int64_t n = f.np;
    if (n == 0) return hipSuccess;
    const int nb = (int)std::min<int64_t>(RED_BLOCKS, grid_for(n, EW_BLOCK));
    bbox_partial_kernel<<<nb, EW_BLOCK, 0, st>>>(f.X, n, f.cap, s.part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    morton_kernel<<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(f.X, n, f.cap, s.part, nb, s.kin,
                                                             s.vin);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = s.sort_tmp_bytes;
    return rocprim::radix_sort_pairs(s.sort_tmp, bytes, s.kin, s.kout, s.vin, s.perm,
                                     (unsigned int)n, 0, 30, st);
}

// Bounds of each 256-slot tile of the sorted field: bbox centre, half
// diagonal, largest sigma.  One workgroup per tile.
__global__ __launch_bounds__(UJ_BLOCK) void tile_bounds_kernel(const double4 *__restrict__ spos,
                                                               const int32_t *__restrict__ perm,
                                                               const double *__restrict__ sg,
                                                               int64_t n, double4 *__restrict__ tb)
{
    __shared__ double red[7][UJ_BLOCK];
    const int t = threadIdx.x;
    const int64_t s = (int64_t)blockIdx.x * UJ_BLOCK + t;
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    if (s < n) {
        const double4 p = spos[s];
        v[0] = v[3] = p.x;
        v[1] = v[4] = p.y;
        v[2] = v[5] = p.z;
        v[6] = sg[perm[s]];
    }
    for (int k = 0; k < 7; ++k) red[k][t] = v[k];
    __syncthreads();
    for (int w = UJ_BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 3; ++k) red[k][t] = fmin(red[k][t], red[k][t + w]);
            for (int k = 3; k < 7; ++k) red[k][t] = fmax(red[k][t], red[k][t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const double hx = 0.5 * (red[3][0] - red[0][0]), hy = 0.5 * (red[4][0] - red[1][0]),
                     hz = 0.5 * (red[5][0] - red[2][0]);
        tb[2 * blockIdx.x] = make_double4(red[0][0] + hx, red[1][0] + hy, red[2][0] + hz,
                                          sqrt(hx * hx + hy * hy + hz * hz));
        tb[2 * blockIdx.x + 1] = make_double4(red[6][0], 0.0, 0.0, 0.0);
    }
}

hipError_t launch_pack_mixed(hipStream_t st, int kernel, const DeviceField &f, Scratch &s)
{
    const int64_t n = f.np;
    if (n == 0) return hipSuccess;
    const unsigned g = grid_for(n, EW_BLOCK);
    switch (kernel) {
    case SINGULAR:
        pack_mixed_kernel<SINGULAR><<<g, EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, s.perm, n, f.cap, s.spos, s.sgam);
        break;
    case GAUSSIAN:
        pack_mixed_kernel<GAUSSIAN><<<g, EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, s.perm, n, f.cap, s.spos, s.sgam);
        break;
    case GAUSSIANERF:
        pack_mixed_kernel<GAUSSIANERF><<<g, EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, s.perm, n, f.cap, s.spos, s.sgam);
        break;
    case WINCKELMANS:
        pack_mixed_kernel<WINCKELMANS><<<g, EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, s.perm, n, f.cap, s.spos, s.sgam);
        break;
    default:
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    tile_bounds_kernel<<<grid_for(n, UJ_BLOCK), UJ_BLOCK, 0, st>>>(s.spos, s.perm, f.sigma, n,
                                                                 s.tbnd);
    return hipGetLastError();
}

hipError_t launch_sort_pack_mixed(hipStream_t st, int kernel, const DeviceField &f, Scratch &s)
{
    hipError_t e = launch_sort_mixed(st, f, s);
    return e != hipSuccess ? e : launch_pack_mixed(st, kernel, f, s);
}

hipError_t launch_uj_mixed(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                           int reset)
{
    if (f.np == 0) return hipSuccess;
    return uj_mixed_any(st, kernel, false, s.spos, f.np, s.perm, s, f.np, f.cap, f.U, f.J, reset,
                        s.tbnd);
}

// Core-size derivatives on the mixed path, from the sorted records and tile
// bounds of the UJ that just ran.
hipError_t launch_dsig_mixed(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                             double *DU, double *DJ)
{
    if (f.np == 0) return hipSuccess;
    switch (kernel) {
    case GAUSSIAN:
        return uj_mixed_go<GAUSSIAN, true>(st, false, s.spos, f.np, s.perm, s.spos, s.sgam, f.np, f.cap, DU, DJ, 1, s.tbnd, s.kpart);
    case GAUSSIANERF:
        return uj_mixed_go<GAUSSIANERF, true>(st, false, s.spos, f.np, s.perm, s.spos, s.sgam, f.np, f.cap, DU, DJ, 1, s.tbnd, s.kpart);
    case WINCKELMANS:
        return uj_mixed_go<WINCKELMANS, true>(st, false, s.spos, f.np, s.perm, s.spos, s.sgam, f.np, f.cap, DU, DJ, 1, s.tbnd, s.kpart);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_uj_mixed_targets(hipStream_t st, int kernel, const DeviceField &f,
                                   const Scratch &s, const double4 *tgt, int64_t nt,
                                   const int32_t *tperm, int64_t ldo, double *U, double *J,
                                   int reset)
{
    if (nt == 0) return hipSuccess;
    return uj_mixed_any(st, kernel, false, tgt, nt, tperm, s, f.np, ldo, U, J, reset, nullptr);
}

// ===========================================================================
// Estr (SFS stretching), FP64
// ===========================================================================

// Source record: (x, y, z, 1/sigma) (Gx, Gy, Gz, 0) (J0..J3) (J4..J7) (J8, 0, 0, 0);
// with the core-size derivatives DJ (dynamic SFS) the last one is
// (J8, DJ0, DJ1, DJ2) followed by (DJ3..DJ6) (DJ7, DJ8, 0, 0): 7 per particle.
template <bool DYN>
__global__ __launch_bounds__(EW_BLOCK) void pack_estr_kernel(const double *__restrict__ X,
                                                             const double *__restrict__ G,
                                                             const double *__restrict__ sg,
                                                             const double *__restrict__ Jm,
                                                             const double *__restrict__ DJ,
                                                             int64_t n, int64_t cap,
                                                             double4 *__restrict__ out)
{
    constexpr int NR = DYN ? 7 : 5;
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    double4 *o = out + NR * i;
    o[0] = make_double4(X[i], X[cap + i], X[2 * cap + i], 1.0 / sg[i]);
    o[1] = make_double4(G[i], G[cap + i], G[2 * cap + i], 0.0);
    o[2] = make_double4(Jm[i], Jm[cap + i], Jm[2 * cap + i], Jm[3 * cap + i]);
    o[3] = make_double4(Jm[4 * cap + i], Jm[5 * cap + i], Jm[6 * cap + i], Jm[7 * cap + i]);
    if constexpr (DYN) {
        o[4] = make_double4(Jm[8 * cap + i], DJ[i], DJ[cap + i], DJ[2 * cap + i]);
        o[5] = make_double4(DJ[3 * cap + i], DJ[4 * cap + i], DJ[5 * cap + i], DJ[6 * cap + i]);
        o[6] = make_double4(DJ[7 * cap + i], DJ[8 * cap + i], 0.0, 0.0);
    } else {
        o[4] = make_double4(Jm[8 * cap + i], 0.0, 0.0, 0.0);
    }
}

// The nine J entries (and with DYN the nine DJ entries) of an Estr record.
template <bool DYN>
__device__ __forceinline__ void estr_rec_j(const double4 *r, double *Jv, double *DJv)
{
    const double4 q2 = r[2], q3 = r[3], q4 = r[4];
    Jv[0] = q2.x; Jv[1] = q2.y; Jv[2] = q2.z; Jv[3] = q2.w;
    Jv[4] = q3.x; Jv[5] = q3.y; Jv[6] = q3.z; Jv[7] = q3.w; Jv[8] = q4.x;
    if constexpr (DYN) {
        const double4 q5 = r[5], q6 = r[6];
        DJv[0] = q4.y; DJv[1] = q4.z; DJv[2] = q4.w;
        DJv[3] = q5.x; DJv[4] = q5.y; DJv[5] = q5.z; DJv[6] = q5.w;
        DJv[7] = q6.x; DJv[8] = q6.y;
    }
}

// S_k = sum_l dJ[l,k] G_l (transposed) or sum_l dJ[k,l] G_l (classic).
template <bool TRANSPOSED>
__device__ __forceinline__ void stretch_of(const double *dJ, double gx, double gy, double gz,
                                           double &S0, double &S1, double &S2)
{
    if constexpr (TRANSPOSED) {
        S0 = dJ[0] * gx + dJ[1] * gy + dJ[2] * gz;
        S1 = dJ[3] * gx + dJ[4] * gy + dJ[5] * gz;
        S2 = dJ[6] * gx + dJ[7] * gy + dJ[8] * gz;
    } else {
        S0 = dJ[0] * gx + dJ[3] * gy + dJ[6] * gz;
        S1 = dJ[1] * gx + dJ[4] * gy + dJ[7] * gz;
        S2 = dJ[2] * gx + dJ[5] * gy + dJ[8] * gz;
    }
}

// Estr of every particle; with DYN also its core-size derivative
//   D[E]_i = sum_j zD_j (J_i - J_j)-stretching + zeta_j (DJ_i - DJ_j)-stretching
// into DE (zeta_zD_f64).
template <int K, bool TRANSPOSED, bool DYN>
__global__ __launch_bounds__(UJ_BLOCK) void estr_f64_kernel(const double4 *__restrict__ src,
                                                            int64_t n, int64_t cap,
                                                            double *__restrict__ E,
                                                            double *__restrict__ DE)
{
    constexpr int NR = DYN ? 7 : 5;
    __shared__ double4 tile[NR * UJ_BLOCK];
    const int t = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * UJ_BLOCK + t;
    const bool active = i < n;
    double xi = 0, yi = 0, zi = 0, Ji[9], DJi[9];
    for (int k = 0; k < 9; ++k) Ji[k] = DJi[k] = 0.0;
    if (active) {
        const double4 p = src[NR * i];
        xi = p.x; yi = p.y; zi = p.z;
        estr_rec_j<DYN>(src + NR * i, Ji, DJi);
    }
    double e0 = 0, e1 = 0, e2 = 0, f0 = 0, f1 = 0, f2 = 0;
    for (int64_t base = 0; base < n; base += UJ_BLOCK) {
        __syncthreads();
        const int64_t js = base + t;
        if (js < n) {
            for (int k = 0; k < NR; ++k) tile[NR * t + k] = src[NR * js + k];
        }
        __syncthreads();
        const int cnt = (int)((n - base) < UJ_BLOCK ? (n - base) : UJ_BLOCK);
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 a = tile[NR * jj], gm = tile[NR * jj + 1];
            double Jj[9], DJj[9];
            estr_rec_j<DYN>(tile + NR * jj, Jj, DJj);
            const double d1 = xi - a.x, d2 = yi - a.y, d3 = zi - a.z;
            const double r = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
            const double is3 = a.w * a.w * a.w;
            double dJ[9], S0, S1, S2;
            for (int k = 0; k < 9; ++k) dJ[k] = Ji[k] - Jj[k];
            stretch_of<TRANSPOSED>(dJ, gm.x, gm.y, gm.z, S0, S1, S2);
            if constexpr (DYN) {
                double z, zD;
                zeta_zD_f64<K>(r * a.w, z, zD);
                const double zs = z * is3, zDs = zD * is3;
                double T0, T1, T2;
                for (int k = 0; k < 9; ++k) dJ[k] = DJi[k] - DJj[k];
                stretch_of<TRANSPOSED>(dJ, gm.x, gm.y, gm.z, T0, T1, T2);
                e0 += zs * S0; e1 += zs * S1; e2 += zs * S2;
                f0 += zDs * S0 + zs * T0;
                f1 += zDs * S1 + zs * T1;
                f2 += zDs * S2 + zs * T2;
            } else {
                const double zs = zeta_f64<K>(r * a.w) * is3;
                e0 += zs * S0; e1 += zs * S1; e2 += zs * S2;
            }
        }
    }
    if (!active) return;
    E[i] = e0;
    E[cap + i] = e1;
    E[2 * cap + i] = e2;
    if constexpr (DYN) {
        DE[i] = f0;
        DE[cap + i] = f1;
        DE[2 * cap + i] = f2;
    }
}

hipError_t launch_estr_f64(hipStream_t st, int kernel, int transposed, const DeviceField &f,
                           Scratch &s, const double *DJ, double *DE)
{
    const int64_t n = f.np;
    if (n == 0) return hipSuccess;
    const bool dyn = DJ != nullptr;
    if (dyn)
        pack_estr_kernel<true><<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, f.J,
                                                                           DJ, n, f.cap, s.srcE);
    else
        pack_estr_kernel<false><<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(f.X, f.G, f.sigma, f.J,
                                                                            DJ, n, f.cap, s.srcE);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned g = grid_for(n, UJ_BLOCK);
#define VPMX_ESTR(KK)                                                                        \
    if (dyn && transposed)                                                                   \
        estr_f64_kernel<KK, true, true><<<g, UJ_BLOCK, 0, st>>>(s.srcE, n, f.cap, f.E, DE);  \
    else if (dyn)                                                                            \
        estr_f64_kernel<KK, false, true><<<g, UJ_BLOCK, 0, st>>>(s.srcE, n, f.cap, f.E, DE); \
    else if (transposed)                                                                     \
        estr_f64_kernel<KK, true, false><<<g, UJ_BLOCK, 0, st>>>(s.srcE, n, f.cap, f.E, DE); \
    else                                                                                     \
        estr_f64_kernel<KK, false, false><<<g, UJ_BLOCK, 0, st>>>(s.srcE, n, f.cap, f.E, DE);
    switch (kernel) {
    case GAUSSIAN: VPMX_ESTR(GAUSSIAN); break;
    case GAUSSIANERF: VPMX_ESTR(GAUSSIANERF); break;
    case WINCKELMANS: VPMX_ESTR(WINCKELMANS); break;
    default: return hipErrorInvalidValue;
    }
#undef VPMX_ESTR
    return hipGetLastError();
}

// C = Cs, clipped to 0 where C Gamma.Estr < 0 (docs/src/theory/rvpm.md:296).
__global__ __launch_bounds__(EW_BLOCK) void sfs_coeff_kernel(int64_t n, int64_t cap,
                                                             const double *__restrict__ G,
                                                             const double *__restrict__ E,
                                                             double *__restrict__ C, double Cs,
                                                             int clip)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    double c = Cs;
    if (clip) {
        const double ge = G[i] * E[i] + G[cap + i] * E[cap + i] + G[2 * cap + i] * E[2 * cap + i];
        if (c * ge < 0.0) c = 0.0;
    }
    C[i] = c;
}

// ---- dynamic SFS procedure (docs/src/theory/rvpm.md:264-296) ---------------
// C_d = <Gamma . L> / <Gamma . m> per particle with the exact core-size
// derivatives DJ = D[J] and DE = D[E] (D = sigma d/dsigma of every core):
//   num = (3 alpha - 2)/sigma  Gamma^T DJ Gamma
//       (L = (3/sigma)(Gamma.grad)(u - ubar) + (Gamma.grad) dubar/dsigma with the
//        unresolved u - ubar taken at the test-filter level alpha sigma to first
//        order in 1 - alpha: u - ubar ~ (alpha - 1) D[ubar]),
//   den = sigma^2/zeta0  Gamma . DE      (m = sigma^3/zeta0 dE/dsigma),
// each averaged along the particle's path as a <- w a_new + (1 - w) a_old (the
// first sample taken as is), C = <num>/<den> bounded to [minC, maxC], then
// clipped to 0 where C Gamma.E < 0.  Static particles keep their C.
__global__ __launch_bounds__(EW_BLOCK) void dyn_coeff_kernel(
    int64_t n, int64_t cap, const double *__restrict__ G, const double *__restrict__ sg,
    const double *__restrict__ E, const double *__restrict__ DJ, const double *__restrict__ DE,
    double *__restrict__ C, const uint8_t *__restrict__ st, double alpha, double zeta0, double w,
    double minC, double maxC, int clip)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (st && st[i]) return;
    const double g[3] = {G[i], G[cap + i], G[2 * cap + i]};
    double q = 0.0;
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) q += g[k] * DJ[(k + 3 * l) * cap + i] * g[l];
    const double sig = sg[i];
    const double nume = (3.0 * alpha - 2.0) / sig * q;
    const double deno = sig * sig / zeta0 *
                        (g[0] * DE[i] + g[1] * DE[cap + i] + g[2] * DE[2 * cap + i]);
    const double n_old = C[cap + i], d_old = C[2 * cap + i];
    double na = nume, da = deno;
    if (n_old != 0.0 || d_old != 0.0) {
        na = w * nume + (1.0 - w) * n_old;
        da = w * deno + (1.0 - w) * d_old;
    }
    C[cap + i] = na;
    C[2 * cap + i] = da;
    double c = da != 0.0 ? na / da : 0.0;
    c = fmin(fmax(c, minC), maxC);
    const double ge = g[0] * E[i] + g[1] * E[cap + i] + g[2] * E[2 * cap + i];
    if (clip && c * ge < 0.0) c = 0.0;
    C[i] = c;
}

hipError_t launch_dyn_coeff(hipStream_t st, const DeviceField &f, double alpha, double zeta0,
                            double rlxf, double minC, double maxC, int clip)
{
    if (f.np == 0) return hipSuccess;
    dyn_coeff_kernel<<<grid_for(f.np, EW_BLOCK), EW_BLOCK, 0, st>>>(
        f.np, f.cap, f.G, f.sigma, f.E, f.DJ(), f.DE(), f.C,
        f.any_static ? f.is_static : nullptr, alpha, zeta0, rlxf, minC, maxC, clip);
    return hipGetLastError();
}

hipError_t launch_sfs_coeff(hipStream_t st, const DeviceField &f, double Cs, int clip)
{
    if (f.np == 0) return hipSuccess;
    sfs_coeff_kernel<<<grid_for(f.np, EW_BLOCK), EW_BLOCK, 0, st>>>(f.np, f.cap, f.G, f.E, f.C, Cs,
                                                                    clip);
    return hipGetLastError();
}

// ===========================================================================
// Streaming passes (FP64)
// ===========================================================================

__global__ __launch_bounds__(EW_BLOCK) void fill_kernel(double *__restrict__ p, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill(hipStream_t st, double *p, int64_t n, double v)
{
    if (n == 0) return hipSuccess;
    fill_kernel<<<grid_for(n, EW_BLOCK), EW_BLOCK, 0, st>>>(p, n, v);
    return hipGetLastError();
}

// One low-storage RK substep (Euler is a = 0, b = 1) of every non-static
// particle:  q <- a q + dt F(y);  y <- y + b q  for y = (X, Gamma, sigma), with
// F the rVPM right-hand side
//   Z     = [ (g+f)/(1+3f) S.G - f/(1+3f) (C sigma^3/zeta0) E.G ] / |G|^2
//   dX/dt = U + Uinf,  dG/dt = S - 3 Z G - (C sigma^3/zeta0) E,  dsigma/dt = -sigma Z.
struct UpdateArgs {
    double a, b, dt, f, g, zeta0, Ux, Uy, Uz;
    int transposed, sfs;
    int viscous;      // 1 core spreading: dsigma/dt += nu / sigma; 2 PSE: dGamma/dt += V
    double nu;
};

__global__ __launch_bounds__(EW_BLOCK) void update_kernel(int64_t n, int64_t cap, UpdateArgs q,
                                                          double *__restrict__ X,
                                                          double *__restrict__ G,
                                                          double *__restrict__ sg,
                                                          const double *__restrict__ U,
                                                          const double *__restrict__ J,
                                                          const double *__restrict__ E,
                                                          const double *__restrict__ C,
                                                          double *__restrict__ M,
                                                          const uint8_t *__restrict__ st,
                                                          const double *__restrict__ uinf,
                                                          const int32_t *__restrict__ fid,
                                                          const double *__restrict__ V)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (st && st[i]) return;
    double Ux = q.Ux, Uy = q.Uy, Uz = q.Uz;
    if (uinf) {
        const int64_t fk = fid ? 3 * (int64_t)fid[i] : 0;
        Ux = uinf[fk];
        Uy = uinf[fk + 1];
        Uz = uinf[fk + 2];
    }
    double Jm[9];
    for (int k = 0; k < 9; ++k) Jm[k] = J[k * cap + i];
    const double G0 = G[i], G1 = G[cap + i], G2 = G[2 * cap + i];
    const double sig = sg[i];
    double S0, S1, S2;
    if (q.transposed) {
        S0 = Jm[0] * G0 + Jm[1] * G1 + Jm[2] * G2;
        S1 = Jm[3] * G0 + Jm[4] * G1 + Jm[5] * G2;
        S2 = Jm[6] * G0 + Jm[7] * G1 + Jm[8] * G2;
    } else {
        S0 = Jm[0] * G0 + Jm[3] * G1 + Jm[6] * G2;
        S1 = Jm[1] * G0 + Jm[4] * G1 + Jm[7] * G2;
        S2 = Jm[2] * G0 + Jm[5] * G1 + Jm[8] * G2;
    }
    double e0 = 0, e1 = 0, e2 = 0, csfs = 0;
    if (q.sfs) {
        e0 = E[i]; e1 = E[cap + i]; e2 = E[2 * cap + i];
        csfs = C[i] * sig * sig * sig / q.zeta0;
    }
    const double Gm2 = G0 * G0 + G1 * G1 + G2 * G2;
    const double SG = S0 * G0 + S1 * G1 + S2 * G2;
    const double EG = e0 * G0 + e1 * G1 + e2 * G2;
    double Z = 0.0;
    if (Gm2 > 0.0)
        Z = ((q.g + q.f) / (1.0 + 3.0 * q.f) * SG - q.f / (1.0 + 3.0 * q.f) * csfs * EG) / Gm2;
    double F[7] = {U[i] + Ux, U[cap + i] + Uy, U[2 * cap + i] + Uz,
                   S0 - 3.0 * Z * G0 - csfs * e0, S1 - 3.0 * Z * G1 - csfs * e1,
                   S2 - 3.0 * Z * G2 - csfs * e2, -sig * Z};
    if (q.viscous == 2) {            // particle strength exchange
        F[3] += V[i];
        F[4] += V[cap + i];
        F[5] += V[2 * cap + i];
    } else if (q.viscous == 1) {     // core spreading
        F[6] += q.nu / sig;
    }
    double Mq[7];
    for (int k = 0; k < 7; ++k) Mq[k] = q.a * M[k * cap + i] + q.dt * F[k];
    for (int k = 0; k < 7; ++k) M[k * cap + i] = Mq[k];
    X[i] += q.b * Mq[0];
    X[cap + i] += q.b * Mq[1];
    X[2 * cap + i] += q.b * Mq[2];
    G[i] = G0 + q.b * Mq[3];
    G[cap + i] = G1 + q.b * Mq[4];
    G[2 * cap + i] = G2 + q.b * Mq[5];
    sg[i] = sig + q.b * Mq[6];
}

hipError_t launch_update(hipStream_t st, const DeviceField &f, int transposed, int sfs,
                         double fcoef, double gcoef, double zeta0, const double Uinf[3],
                         double a, double b, double dt)
{
    if (f.np == 0) return hipSuccess;
    UpdateArgs q{a, b, dt, fcoef, gcoef, zeta0, Uinf[0], Uinf[1], Uinf[2], transposed, sfs, 0, 0.0};
    update_kernel<<<grid_for(f.np, EW_BLOCK), EW_BLOCK, 0, st>>>(
        f.np, f.cap, q, f.X, f.G, f.sigma, f.U, f.J, f.E, f.C, f.M,
        f.any_static ? f.is_static : nullptr, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_update_dev(hipStream_t st, const DeviceField &f, int transposed, int sfs,
                             double fcoef, double gcoef, double zeta0, const double *uinf,
                             const int32_t *fid, double a, double b, double dt, int viscous,
                             double nu)
{
    if (f.np == 0) return hipSuccess;
    if (viscous == 2 && !f.V) return hipErrorInvalidValue;
    UpdateArgs q{a, b, dt, fcoef, gcoef, zeta0, 0.0, 0.0, 0.0, transposed, sfs, viscous, nu};
    update_kernel<<<grid_for(f.np, EW_BLOCK), EW_BLOCK, 0, st>>>(
        f.np, f.cap, q, f.X, f.G, f.sigma, f.U, f.J, f.E, f.C, f.M,
        f.any_static ? f.is_static : nullptr, uinf, fid, f.V);
    return hipGetLastError();
}

// Pedrizzetti relaxation; kind 2 rescales to keep |Gamma|.
__global__ __launch_bounds__(EW_BLOCK) void relax_kernel(int64_t n, int64_t cap, int kind,
                                                         double rlxf, const double *__restrict__ J,
                                                         double *__restrict__ G,
                                                         const uint8_t *__restrict__ st)
{
    const int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (st && st[i]) return;
    const double w1 = J[5 * cap + i] - J[7 * cap + i];
    const double w2 = J[6 * cap + i] - J[2 * cap + i];
    const double w3 = J[cap + i] - J[3 * cap + i];
    const double nw = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
    if (nw == 0.0) return;
    double g0 = G[i], g1 = G[cap + i], g2 = G[2 * cap + i];
    const double nG = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    const double cosang = (g0 * w1 + g1 * w2 + g2 * w3) / (nG * nw);
    g0 = (1.0 - rlxf) * g0 + rlxf * nG * w1 / nw;
    g1 = (1.0 - rlxf) * g1 + rlxf * nG * w2 / nw;
    g2 = (1.0 - rlxf) * g2 + rlxf * nG * w3 / nw;
    if (kind == 2) {
        const double bb = sqrt(1.0 - 2.0 * (1.0 - rlxf) * rlxf * (1.0 - cosang));
        g0 /= bb; g1 /= bb; g2 /= bb;
    }
    G[i] = g0;
    G[cap + i] = g1;
    G[2 * cap + i] = g2;
}

hipError_t launch_relax(hipStream_t st, const DeviceField &f, int kind, double rlxf)
{
    if (f.np == 0 || kind == 0) return hipSuccess;
    relax_kernel<<<grid_for(f.np, EW_BLOCK), EW_BLOCK, 0, st>>>(
        f.np, f.cap, kind, rlxf, f.J, f.G, f.any_static ? f.is_static : nullptr);
    return hipGetLastError();
}

// Enstrophy and energy: per-block partials, then one block sums them in a
// fixed order (deterministic).
__global__ __launch_bounds__(EW_BLOCK) void diag_partial_kernel(int64_t n, int64_t cap,
                                                                const double *__restrict__ X,
                                                                const double *__restrict__ G,
                                                                const double *__restrict__ U,
                                                                const double *__restrict__ J,
                                                                double *__restrict__ part)
{
    __shared__ double r0[EW_BLOCK], r1[EW_BLOCK];
    double ens = 0.0, en = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * EW_BLOCK + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * EW_BLOCK) {
        const double g0 = G[i], g1 = G[cap + i], g2 = G[2 * cap + i];
        const double w1 = J[5 * cap + i] - J[7 * cap + i];
        const double w2 = J[6 * cap + i] - J[2 * cap + i];
        const double w3 = J[cap + i] - J[3 * cap + i];
        ens += g0 * w1 + g1 * w2 + g2 * w3;
        const double x0 = X[i], x1 = X[cap + i], x2 = X[2 * cap + i];
        const double c0 = x1 * g2 - x2 * g1, c1 = x2 * g0 - x0 * g2, c2 = x0 * g1 - x1 * g0;
        en += U[i] * c0 + U[cap + i] * c1 + U[2 * cap + i] * c2;
    }
    r0[threadIdx.x] = ens;
    r1[threadIdx.x] = en;
    __syncthreads();
    for (int w = EW_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r0[0];
        part[2 * blockIdx.x + 1] = r1[0];
    }
}

__global__ void diag_final_kernel(const double *__restrict__ part, int nparts,
                                  double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nparts; ++k) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    out[0] = 0.5 * a;
    out[1] = b;
}

hipError_t launch_diagnostics(hipStream_t st, const DeviceField &f, Scratch &s, double *out2)
{
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, grid_for(f.np, EW_BLOCK)));
    diag_partial_kernel<<<nb, EW_BLOCK, 0, st>>>(f.np, f.cap, f.X, f.G, f.U, f.J, s.part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    diag_final_kernel<<<1, 64, 0, st>>>(s.part, nb, out2);
    return hipGetLastError();
}

double zeta0_of(int kernel)
{
    switch (kernel) {
    case GAUSSIAN: return 0.238732414637843;        // 3/(4 pi)
    case GAUSSIANERF: return 0.06349363593424097;   // (2 pi)^-3/2
    case WINCKELMANS: return 0.5968310365946075;    // 15/(8 pi)
    default: return 1.0;
    }
}

}  // namespace vpmx
