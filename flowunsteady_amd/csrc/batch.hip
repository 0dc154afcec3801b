// batch.hip -- many independent particle fields of uneven size resident as one
// concatenation, and their pairwise pass in a single launch.
//
// Field k owns slots [off[k], off[k+1]).  The mixed path sorts every field by
// its own 30-bit Morton key (computed in the field's own bounding box) with the
// field index in the high bits, so the sorted fields stay contiguous at the same
// offsets and each is ordered exactly as it would be alone.  The pairwise pass
// walks a work list of (target block, source range) items, one per block of
// `BLK` consecutive targets of one field, ordered largest field first (longest
// processing time first), either as one workgroup per item (flat) or through a
// device work queue drained by a persistent grid sized to the resident
// occupancy (queue).  Fields never interact.  Sweeps of independent cases
// (each advanced by its own `nextstep`, src/FLOWUnsteady_simulation.jl:358)
// become one launch instead of one under-filled launch per case.
#include <rocprim/device/device_radix_sort.hpp>

#include "device_math.h"
#include "internal.h"
#include "uj_block.h"

namespace vpmx {

namespace {

constexpr int EWB = 256;

inline unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// fid[i] = k with off[k] <= i < off[k+1] (binary search over nf+1 offsets).
__global__ __launch_bounds__(EWB) void fid_kernel(const int64_t *__restrict__ off, int nf,
                                                  int64_t n, int32_t *__restrict__ fid)
{
    const int64_t i = (int64_t)blockIdx.x * EWB + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nf;   // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    fid[i] = lo;
}

// One workgroup per field: lo xyz, hi xyz of the field's positions.
__global__ __launch_bounds__(EWB) void field_bbox_kernel(const double *__restrict__ X,
                                                         int64_t cap,
                                                         const int64_t *__restrict__ off,
                                                         double *__restrict__ box)
{
    __shared__ double red[6][EWB];
    const int f = blockIdx.x;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = off[f] + threadIdx.x; i < off[f + 1]; i += EWB) {
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
    for (int w = EWB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
                red[3 + k][threadIdx.x] = fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) box[6 * f + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(EWB) void field_morton_kernel(const double *__restrict__ X,
                                                           int64_t n, int64_t cap,
                                                           const int32_t *__restrict__ fid,
                                                           const double *__restrict__ box,
                                                           uint64_t *__restrict__ key,
                                                           int32_t *__restrict__ val)
{
    const int64_t i = (int64_t)blockIdx.x * EWB + threadIdx.x;
    if (i >= n) return;
    const int f = fid[i];
    const uint32_t m = morton30(box + 6 * f, X[i], X[cap + i], X[2 * cap + i]);
    key[i] = ((uint64_t)f << 30) | m;
    val[i] = (int32_t)i;
}

__global__ void queue_reset_kernel(int32_t *head)
{
    if (threadIdx.x == 0) head[0] = 0;
}

// ---- pairwise pass over the work list -----------------------------------------
template <int K, int BLK>
__global__ __launch_bounds__(BLK) void uj_mixed_items_kernel(const BatchItem *__restrict__ items,
                                                             const int32_t *__restrict__ perm,
                                                             const double4 *__restrict__ spos,
                                                             const float4 *__restrict__ sgam,
                                                             const double4 *__restrict__ tbnd,
                                                             int64_t ldo, double *__restrict__ U,
                                                             double *__restrict__ J, int reset)
{
    __shared__ float4 tpos[BLK];
    __shared__ float4 tgam[BLK];
    __shared__ float4 tlo[BLK];
    const BatchItem it = items[blockIdx.x];
    uj_mixed_block<K, BLK>(spos, it.t0, it.t1, perm, spos, sgam, it.s0, it.s1, ldo, U, J, reset,
                           tpos, tgam, tlo, tbnd + 2 * it.tb, it.s0);
}

// Persistent grid: each workgroup takes the next item from the queue head until
// the list is exhausted.  The item index is broadcast through LDS so the whole
// workgroup leaves the loop together.
template <int K, int BLK>
__global__ __launch_bounds__(BLK) void uj_mixed_queue_kernel(const BatchItem *__restrict__ items,
                                                             int nitems, int32_t *head,
                                                             const int32_t *__restrict__ perm,
                                                             const double4 *__restrict__ spos,
                                                             const float4 *__restrict__ sgam,
                                                             const double4 *__restrict__ tbnd,
                                                             int64_t ldo, double *__restrict__ U,
                                                             double *__restrict__ J, int reset)
{
    __shared__ float4 tpos[BLK];
    __shared__ float4 tgam[BLK];
    __shared__ float4 tlo[BLK];
    __shared__ int32_t cur;
    for (;;) {
        if (threadIdx.x == 0) cur = atomicAdd(head, 1);
        __syncthreads();
        const int k = cur;
        __syncthreads();
        if (k >= nitems) break;
        const BatchItem it = items[k];
        uj_mixed_block<K, BLK>(spos, it.t0, it.t1, perm, spos, sgam, it.s0, it.s1, ldo, U, J,
                               reset, tpos, tgam, tlo, tbnd + 2 * it.tb, it.s0);
    }
}

template <int K, int BLK>
__global__ __launch_bounds__(BLK) void uj_f64_items_kernel(const BatchItem *__restrict__ items,
                                                           const double4 *__restrict__ src,
                                                           const double *__restrict__ X,
                                                           int64_t cap, double *__restrict__ U,
                                                           double *__restrict__ J, int reset)
{
    __shared__ double4 tile[2 * BLK];
    const BatchItem it = items[blockIdx.x];
    uj_f64_block<K, BLK>(src, it.s0, it.s1, X, it.t0, it.t1, cap, U, J, reset, tile);
}

template <int K, int BLK>
hipError_t mixed_go(hipStream_t st, const DeviceField &f, const Scratch &s, BatchState &b,
                    int reset, int sched, int grid)
{
    const int n = (int)b.nitems;
    if (sched == SCHED_QUEUE) {
        queue_reset_kernel<<<1, 64, 0, st>>>(b.counter);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int g = std::max(1, std::min(grid, n));
        uj_mixed_queue_kernel<K, BLK><<<g, BLK, 0, st>>>(b.items, n, b.counter, s.perm, s.spos,
                                                         s.sgam, b.tbnd, f.cap, f.U, f.J, reset);
    } else {
        uj_mixed_items_kernel<K, BLK><<<n, BLK, 0, st>>>(b.items, s.perm, s.spos, s.sgam, b.tbnd,
                                                         f.cap, f.U, f.J, reset);
    }
    return hipGetLastError();
}

template <int BLK>
hipError_t mixed_k(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                   BatchState &b, int reset, int sched, int grid)
{
    switch (kernel) {
    case SINGULAR: return mixed_go<SINGULAR, BLK>(st, f, s, b, reset, sched, grid);
    case GAUSSIAN: return mixed_go<GAUSSIAN, BLK>(st, f, s, b, reset, sched, grid);
    case GAUSSIANERF: return mixed_go<GAUSSIANERF, BLK>(st, f, s, b, reset, sched, grid);
    case WINCKELMANS: return mixed_go<WINCKELMANS, BLK>(st, f, s, b, reset, sched, grid);
    default: return hipErrorInvalidValue;
    }
}

template <int K, int BLK>
hipError_t f64_go(hipStream_t st, const DeviceField &f, const Scratch &s, BatchState &b, int reset)
{
    uj_f64_items_kernel<K, BLK><<<(unsigned)b.nitems, BLK, 0, st>>>(b.items, s.src64, f.X, f.cap,
                                                                    f.U, f.J, reset);
    return hipGetLastError();
}

template <int BLK>
hipError_t f64_k(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                 BatchState &b, int reset)
{
    switch (kernel) {
    case SINGULAR: return f64_go<SINGULAR, BLK>(st, f, s, b, reset);
    case GAUSSIAN: return f64_go<GAUSSIAN, BLK>(st, f, s, b, reset);
    case GAUSSIANERF: return f64_go<GAUSSIANERF, BLK>(st, f, s, b, reset);
    case WINCKELMANS: return f64_go<WINCKELMANS, BLK>(st, f, s, b, reset);
    default: return hipErrorInvalidValue;
    }
}

template <int K, int BLK>
int occ_of()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, uj_mixed_queue_kernel<K, BLK>, BLK, 0) !=
        hipSuccess)
        return 0;
    return nb;
}

template <int BLK>
int occ_k(int kernel)
{
    switch (kernel) {
    case SINGULAR: return occ_of<SINGULAR, BLK>();
    case GAUSSIAN: return occ_of<GAUSSIAN, BLK>();
    case GAUSSIANERF: return occ_of<GAUSSIANERF, BLK>();
    case WINCKELMANS: return occ_of<WINCKELMANS, BLK>();
    default: return 0;
    }
}

}  // namespace

size_t batch_sort_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, bytes, (const uint64_t *)nullptr,
                                    (uint64_t *)nullptr, (const int32_t *)nullptr,
                                    (int32_t *)nullptr, (unsigned int)n, 0, 62, (hipStream_t)0);
    return bytes;
}

hipError_t launch_batch_fid(hipStream_t st, const BatchState &b, int64_t n)
{
    if (n == 0) return hipSuccess;
    fid_kernel<<<grid_of(n, EWB), EWB, 0, st>>>(b.doff, b.nfields, n, b.fid);
    return hipGetLastError();
}

hipError_t launch_batch_sort(hipStream_t st, const DeviceField &f, Scratch &s, const BatchState &b)
{
    const int64_t n = f.np;
    if (n == 0) return hipSuccess;
    field_bbox_kernel<<<b.nfields, EWB, 0, st>>>(f.X, f.cap, b.doff, b.box);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    field_morton_kernel<<<grid_of(n, EWB), EWB, 0, st>>>(f.X, n, f.cap, b.fid, b.box, b.kin,
                                                         s.vin);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    int fbits = 0;
    while ((1 << fbits) < b.nfields) ++fbits;
    size_t bytes = b.sort_tmp_bytes;
    return rocprim::radix_sort_pairs(b.sort_tmp, bytes, b.kin, b.kout, s.vin, s.perm,
                                     (unsigned int)n, 0, 30 + fbits, st);
}

// Bounds of listed slot ranges (each at most 256 slots) of the sorted
// concatenation: bbox centre, half diagonal, largest sigma -- the per-field
// tiles of the batched pass.  One workgroup per tile.
__global__ __launch_bounds__(256) void tile_bounds_list_kernel(const double4 *__restrict__ spos,
                                                               const int32_t *__restrict__ perm,
                                                               const double *__restrict__ sg,
                                                               const int2 *__restrict__ tiles,
                                                               double4 *__restrict__ tb)
{
    __shared__ double red[7][256];
    const int t = threadIdx.x;
    const int2 T = tiles[blockIdx.x];
    const int s = T.x + t;
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    if (s < T.y) {
        const double4 p = spos[s];
        v[0] = v[3] = p.x;
        v[1] = v[4] = p.y;
        v[2] = v[5] = p.z;
        v[6] = sg[perm[s]];
    }
    for (int k = 0; k < 7; ++k) red[k][t] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
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

hipError_t launch_batch_tile_bounds(hipStream_t st, const DeviceField &f, const Scratch &s,
                                    const BatchState &b)
{
    if (b.ntiles == 0) return hipSuccess;
    tile_bounds_list_kernel<<<(unsigned)b.ntiles, 256, 0, st>>>(s.spos, s.perm, f.sigma, b.tiles,
                                                               b.tbnd);
    return hipGetLastError();
}

hipError_t launch_batch_uj_mixed(hipStream_t st, int kernel, const DeviceField &f,
                                 const Scratch &s, BatchState &b, int reset, int sched, int block,
                                 int grid)
{
    if (b.nitems == 0) return hipSuccess;
    if (block != b.items_block) return hipErrorInvalidValue;
    switch (block) {
    case 64: return mixed_k<64>(st, kernel, f, s, b, reset, sched, grid);
    case 128: return mixed_k<128>(st, kernel, f, s, b, reset, sched, grid);
    case 256: return mixed_k<256>(st, kernel, f, s, b, reset, sched, grid);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_uj_f64(hipStream_t st, int kernel, const DeviceField &f,
                               const Scratch &s, BatchState &b, int reset)
{
    if (b.nitems == 0) return hipSuccess;
    switch (b.items_block) {
    case 64: return f64_k<64>(st, kernel, f, s, b, reset);
    case 128: return f64_k<128>(st, kernel, f, s, b, reset);
    case 256: return f64_k<256>(st, kernel, f, s, b, reset);
    default: return hipErrorInvalidValue;
    }
}

int batch_queue_occupancy(int kernel, int block)
{
    switch (block) {
    case 64: return occ_k<64>(kernel);
    case 128: return occ_k<128>(kernel);
    case 256: return occ_k<256>(kernel);
    default: return 0;
    }
}

}  // namespace vpmx
