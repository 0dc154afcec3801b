// fmm.hip -- fast multipole UJ (the UJ_fmm scheme) for gfx950.
//
// Same outputs as the direct UJ (U and J of every particle from every
// particle) in O(N) work, controlled by the expansion order p, the leaf size
// ncrit and the acceptance parameter theta (vpm.FMM(; p, ncrit, theta),
// src/FLOWUnsteady_simulation.jl:43,134).
//
// Far field: the vector potential psi = (1/4 pi) sum_j Gamma_j / |x - x_j| as
// three Laplace potentials, Cartesian Taylor expansions truncated at total
// degree p (raw multipole moments m_a = sum q (x - c)^a, locals
// phi = sum l_b (x - a)^b, b_k = d^k(1/r)/k! by the recurrence
// R^2 n b_k = -(2n-1) sum_i r_i b_{k-e_i} - (n-1) sum_i b_{k-2e_i}), then
// U = curl psi and J = grad U from the first and second derivatives of the
// local expansions.  Near field: the regularised pair kernel (FP64 or mixed),
// exactly as the direct UJ.
//
// Tree: adaptive octree over 63-bit Morton keys of the bounding cube, leaves of
// at most ncrit particles, breadth-first cell order.  Traversal: dual tree,
// (target A, source B) goes to M2L when (R_A + R_B) < theta |c_A - c_B| and the
// gap |c_A - c_B| - R_A - R_B >= sigcut * max sigma in B (so the singular far
// field matches the regularised kernel to its tail), else the larger cell is
// split; leaf-leaf pairs go to P2P.  The tree build and traversal run on the
// host from the GPU-sorted keys; every expansion and pair kernel runs on the
// GPU.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <utility>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "device_math.h"
#include "fmm.h"
#include "internal.h"

namespace vpmx {

constexpr int KEYBITS = 21;
constexpr int FMM_BLOCK = 256;
constexpr int P2P_BLOCK = 64;

static inline unsigned gridn(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// Multi-index tables (host)
// ---------------------------------------------------------------------------
struct MIndex {
    int P = -1, K = 0;
    std::vector<int> mi;     // 3 per index, graded order
    std::vector<int> lut;    // (a*(P+1)+b)*(P+1)+c -> index or -1
    void build(int P_)
    {
        P = P_;
        K = (P + 1) * (P + 2) * (P + 3) / 6;
        mi.assign(3 * K, 0);
        lut.assign((P + 1) * (P + 1) * (P + 1), -1);
        int k = 0;
        for (int n = 0; n <= P; n++)
            for (int a = n; a >= 0; a--)
                for (int b = n - a; b >= 0; b--) {
                    int c = n - a - b;
                    mi[3 * k] = a; mi[3 * k + 1] = b; mi[3 * k + 2] = c;
                    lut[(a * (P + 1) + b) * (P + 1) + c] = k;
                    k++;
                }
    }
    int idx(int a, int b, int c) const
    {
        if (a < 0 || b < 0 || c < 0 || a + b + c > P) return -1;
        return lut[(a * (P + 1) + b) * (P + 1) + c];
    }
};

static double binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

__global__ __launch_bounds__(FMM_BLOCK) void morton63_kernel(const double *__restrict__ X, int64_t n,
                                                             int64_t cap, double lx, double ly,
                                                             double lz, double scale,
                                                             uint64_t *__restrict__ key,
                                                             int32_t *__restrict__ val)
{
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i >= n) return;
    const double lo[3] = {lx, ly, lz};
    uint64_t q[3];
    for (int k = 0; k < 3; k++) {
        double v = (X[k * cap + i] - lo[k]) * scale;
        v = fmin(fmax(v, 0.0), 2097151.0);
        q[k] = (uint64_t)v;
    }
    key[i] = spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
    val[i] = (int32_t)i;
}

// Sorted copies: FP64 source records (x, y, z, Gx) (Gy, Gz, sigma, 1/sigma)
// and sigma in sorted order.
__global__ __launch_bounds__(FMM_BLOCK) void fmm_gather_kernel(const double *__restrict__ X,
                                                               const double *__restrict__ G,
                                                               const double *__restrict__ sg,
                                                               const int32_t *__restrict__ perm,
                                                               int64_t n, int64_t cap,
                                                               double4 *__restrict__ rec,
                                                               double *__restrict__ ssig)
{
    const int64_t s = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm[s];
    const double sig = sg[i];
    rec[2 * s] = make_double4(X[i], X[cap + i], X[2 * cap + i], G[i]);
    rec[2 * s + 1] = make_double4(G[cap + i], G[2 * cap + i], sig, 1.0 / sig);
    ssig[s] = sig;
}

// P2M: one block per leaf; thread k accumulates m_k for the three components.
__global__ __launch_bounds__(FMM_BLOCK) void p2m_kernel(const int32_t *__restrict__ leaves,
                                                        const FmmCellDev *__restrict__ cells,
                                                        const double4 *__restrict__ rec, int K,
                                                        const int *__restrict__ mi,
                                                        double *__restrict__ M)
{
    __shared__ double tp[FMM_BLOCK][3];
    __shared__ double gp[FMM_BLOCK][3];
    const int c = leaves[blockIdx.x];
    const FmmCellDev C = cells[c];
    double acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int base = C.begin; base < C.end; base += FMM_BLOCK) {
        __syncthreads();
        const int j = base + threadIdx.x;
        if (j < C.end) {
            const double4 a = rec[2 * j], b = rec[2 * j + 1];
            tp[threadIdx.x][0] = a.x - C.cx;
            tp[threadIdx.x][1] = a.y - C.cy;
            tp[threadIdx.x][2] = a.z - C.cz;
            gp[threadIdx.x][0] = a.w;
            gp[threadIdx.x][1] = b.x;
            gp[threadIdx.x][2] = b.y;
        }
        __syncthreads();
        const int cnt = min(FMM_BLOCK, C.end - base);
        for (int r = 0; r < 2; r++) {
            const int k = threadIdx.x + r * FMM_BLOCK;
            if (k >= K) continue;
            const int e0 = mi[3 * k], e1 = mi[3 * k + 1], e2 = mi[3 * k + 2];
            for (int jj = 0; jj < cnt; jj++) {
                double px = 1.0, py = 1.0, pz = 1.0;
                for (int e = 0; e < e0; e++) px *= tp[jj][0];
                for (int e = 0; e < e1; e++) py *= tp[jj][1];
                for (int e = 0; e < e2; e++) pz *= tp[jj][2];
                const double mono = px * py * pz;
                acc[r][0] += gp[jj][0] * mono;
                acc[r][1] += gp[jj][1] * mono;
                acc[r][2] += gp[jj][2] * mono;
            }
        }
    }
    for (int r = 0; r < 2; r++) {
        const int k = threadIdx.x + r * FMM_BLOCK;
        if (k >= K) continue;
        for (int q = 0; q < 3; q++) M[((size_t)c * 3 + q) * K + k] = acc[r][q];
    }
}

// monomials d^a for |a| <= P into LDS (thread-parallel over a)
__device__ __forceinline__ void lds_monomials(const double d[3], int K, const int *__restrict__ mi,
                                              double *out)
{
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        double px = 1.0, py = 1.0, pz = 1.0;
        for (int e = 0; e < mi[3 * k]; e++) px *= d[0];
        for (int e = 0; e < mi[3 * k + 1]; e++) py *= d[1];
        for (int e = 0; e < mi[3 * k + 2]; e++) pz *= d[2];
        out[k] = px * py * pz;
    }
}

// M2M for the parents listed in `parents` (one level): m_a(c) += sum_child
// sum_{b <= a} C(a,b) m_b(child) d^{a-b}, d = c_child - c.
__global__ __launch_bounds__(FMM_BLOCK) void m2m_kernel(const int32_t *__restrict__ parents,
                                                        const FmmCellDev *__restrict__ cells,
                                                        int K, const int *__restrict__ mi,
                                                        const int *__restrict__ toff,
                                                        const int2 *__restrict__ tidx,
                                                        const double *__restrict__ tcoef,
                                                        double *__restrict__ M)
{
    extern __shared__ double sh[];
    double *mono = sh;           // K
    double *mc = sh + K;         // 3K child multipole
    const int c = parents[blockIdx.x];
    const FmmCellDev C = cells[c];
    double acc[2][3] = {{0, 0, 0}, {0, 0, 0}};
    for (int ch = 0; ch < C.nchild; ch++) {
        const int d = C.child0 + ch;
        const FmmCellDev D = cells[d];
        __syncthreads();
        const double dd[3] = {D.cx - C.cx, D.cy - C.cy, D.cz - C.cz};
        lds_monomials(dd, K, mi, mono);
        for (int k = threadIdx.x; k < 3 * K; k += blockDim.x) mc[k] = M[(size_t)d * 3 * K + k];
        __syncthreads();
        for (int r = 0; r < 2; r++) {
            const int k = threadIdx.x + r * FMM_BLOCK;
            if (k >= K) continue;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int e = toff[k]; e < toff[k + 1]; e++) {
                const int2 ix = tidx[e];
                const double w = tcoef[e] * mono[ix.y];
                s0 += mc[ix.x] * w;
                s1 += mc[K + ix.x] * w;
                s2 += mc[2 * K + ix.x] * w;
            }
            acc[r][0] += s0;
            acc[r][1] += s1;
            acc[r][2] += s2;
        }
    }
    for (int r = 0; r < 2; r++) {
        const int k = threadIdx.x + r * FMM_BLOCK;
        if (k >= K) continue;
        for (int q = 0; q < 3; q++) M[((size_t)c * 3 + q) * K + k] = acc[r][q];
    }
}

// M2L: one block per target cell, l_b = sum_sources sum_a coef(a, b) m_a D_{a+b},
// D_k the derivatives of 1/|r| at the centre separation.  Sources are taken
// M2L_BATCH at a time: wave w builds the derivative tensor of source w degree
// by degree (the recurrence of fmm_build_tables' rec1/rec2 tables) while the
// block stages the batch's multipoles; then each thread contracts its R chunks
// of the (a, a+b) term list (chunks of one b, evenly sized) against every source
// of the batch.  At the end the chunk sums of each b are added in chunk order:
// the summation order is fixed, so results are bitwise repeatable.  Writes
// (overwrites) the cell's local expansion.
constexpr int M2L_BATCH = FMM_BLOCK / 64;

template <int R>
__global__ __launch_bounds__(FMM_BLOCK) void m2l_kernel(const FmmCellDev *__restrict__ cells,
                                                        const int64_t *__restrict__ off,
                                                        const int32_t *__restrict__ src, int K,
                                                        int P, const int4 *__restrict__ rec1,
                                                        const int4 *__restrict__ rec2,
                                                        const int4 *__restrict__ chunks,
                                                        int nchunk,
                                                        const int *__restrict__ cfirst,
                                                        const int2 *__restrict__ tidx,
                                                        const double *__restrict__ tcoef,
                                                        const double *__restrict__ M,
                                                        double *__restrict__ L)
{
    extern __shared__ double sh[];
    double *bb = sh;                       // M2L_BATCH x K derivatives
    double *ms = sh + M2L_BATCH * K;       // M2L_BATCH x 3K multipoles
    const int c = blockIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const FmmCellDev A = cells[c];
    const int64_t e0 = off[c], e1 = off[c + 1];
    double acc[R][3];
    for (int r = 0; r < R; r++) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    for (int64_t eb = e0; eb < e1; eb += M2L_BATCH) {
        const int nb = (int)min<int64_t>(M2L_BATCH, e1 - eb);
        __syncthreads();
        for (int k = threadIdx.x; k < M2L_BATCH * 3 * K; k += FMM_BLOCK) {
            const int q = k / (3 * K);
            ms[k] = q < nb ? M[(size_t)src[eb + q] * 3 * K + (k - q * 3 * K)] : 0.0;
        }
        double r0 = 0.0, r1 = 0.0, r2 = 0.0, R2 = 1.0;
        double *bw = bb + w * K;
        if (w < nb) {
            const FmmCellDev B = cells[src[eb + w]];
            r0 = A.cx - B.cx; r1 = A.cy - B.cy; r2 = A.cz - B.cz;
            R2 = r0 * r0 + r1 * r1 + r2 * r2;
            if (lane == 0) bw[0] = 1.0 / sqrt(R2);
        } else {
            for (int k = lane; k < K; k += 64) bw[k] = 0.0;
        }
        __syncthreads();
        int kbeg = 1;
        for (int n = 1; n <= P; n++) {
            const int kend = (n + 1) * (n + 2) * (n + 3) / 6;
            if (w < nb) {
                for (int k = kbeg + lane; k < kend; k += 64) {
                    const int4 j1 = rec1[k];   // idx(k - e_i), -1 if none
                    const int4 j2 = rec2[k];   // idx(k - 2 e_i)
                    double s1 = 0.0, s2 = 0.0;
                    if (j1.x >= 0) s1 += r0 * bw[j1.x];
                    if (j1.y >= 0) s1 += r1 * bw[j1.y];
                    if (j1.z >= 0) s1 += r2 * bw[j1.z];
                    if (j2.x >= 0) s2 += bw[j2.x];
                    if (j2.y >= 0) s2 += bw[j2.y];
                    if (j2.z >= 0) s2 += bw[j2.z];
                    bw[k] = -((2.0 * n - 1.0) * s1 + (n - 1.0) * s2) / (n * R2);
                }
            }
            kbeg = kend;
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int ch = threadIdx.x + r * FMM_BLOCK;
            if (ch >= nchunk) continue;
            const int4 C = chunks[ch];   // (b, first term, end term, -)
            double s[M2L_BATCH][3];
#pragma unroll
            for (int q = 0; q < M2L_BATCH; q++) s[q][0] = s[q][1] = s[q][2] = 0.0;
            for (int t = C.y; t < C.z; t++) {
                const int2 ix = tidx[t];   // (a, a+b)
                const double cf = tcoef[t];
#pragma unroll
                for (int q = 0; q < M2L_BATCH; q++) {
                    const double wq = cf * bb[q * K + ix.y];
                    const double *mq = ms + q * 3 * K;
                    s[q][0] += wq * mq[ix.x];
                    s[q][1] += wq * mq[K + ix.x];
                    s[q][2] += wq * mq[2 * K + ix.x];
                }
            }
#pragma unroll
            for (int q = 0; q < M2L_BATCH; q++) {
                acc[r][0] += s[q][0];
                acc[r][1] += s[q][1];
                acc[r][2] += s[q][2];
            }
        }
    }
    // chunk sums -> l_b in chunk order
    __syncthreads();
    double *part = sh;   // nchunk x 3
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ch = threadIdx.x + r * FMM_BLOCK;
        if (ch < nchunk)
            for (int q = 0; q < 3; q++) part[3 * ch + q] = acc[r][q];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += FMM_BLOCK) {
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        for (int ch = cfirst[k]; ch < cfirst[k + 1]; ch++) {
            v0 += part[3 * ch];
            v1 += part[3 * ch + 1];
            v2 += part[3 * ch + 2];
        }
        L[((size_t)c * 3 + 0) * K + k] = v0;
        L[((size_t)c * 3 + 1) * K + k] = v1;
        L[((size_t)c * 3 + 2) * K + k] = v2;
    }
}

// ---------------------------------------------------------------------------
// M2L on the matrix cores
// ---------------------------------------------------------------------------
// With E_k = d^k (1/|r|) (the raw derivatives, k! times the Taylor
// coefficients), m'_a = (-1)^|a| m_a / a! and l'_b = b! l_b the translation is
//     l'_b = sum_a E_{a+b} m'_a,   |a| + |b| <= P,
// a K x K matrix (a Hankel matrix in the multi-index) times the 3 columns of
// the source's multipole.  One wave owns one target cell and takes its sources
// four at a time, one per block of v_mfma_f64_4x4x4f64 (four independent
// 4 x 4 x 4 products): the A operand of block j is a tile of source j's matrix,
// gathered from its E in LDS through the tile index table; B is the source's
// m' (columns q = 0, 1, 2 and a zero fourth column); C accumulates l' of the
// target over every source of the block, one 4 x 4 tile per b-tile, in
// registers.  At the end the four blocks are added (fixed order) and scaled
// by 1/b!.
//
// Operand layouts of v_mfma_f64_4x4x4f64 (lane of element (row, k) of A,
// (k, n) of B and (row, n) of C in block j): A at row + 4 j + 16 k, B at
// n + 4 j + 16 k, C at n + 4 j + 16 row.
//
// E of a source from the Taylor coefficients t_k = E_k / k! of 1/r, whose
// recurrence (that of E_k divided by k!; the k_i factors cancel)
//     n R^2 t_k = -(2n - 1) sum_i r_i t_{k - e_i} - (n - 1) sum_i t_{k - 2 e_i},
// n = |k|, runs degree by degree, 16 lanes per source, in place; a last pass
// multiplies by k!.
// Tile t -> (ta, tb), in the order the tile index table holds them.
template <int P>
struct M2LTiles {
    struct Map {
        int ta[m2l_ntiles8(P)], tb[m2l_ntiles8(P)];
    };
    static constexpr Map make()
    {
        Map m{};
        int nat[m2l_ntile1(P)] = {};
        for (int tb = 0; tb < m2l_ntile1(P); ++tb) nat[tb] = m2l_natile(P, tb);
        int t = 0;
        for (int ta = 0; ta < m2l_ntile1(P); ++ta)
            for (int tb = 0; tb < m2l_ntile1(P) && nat[tb] > ta; ++tb, ++t) {
                m.ta[t] = ta;
                m.tb[t] = tb;
            }
        return m;
    }
    static constexpr Map map = make();
};

template <int P, int T>
__device__ __forceinline__ void m2l_mfma_one(double (&C)[m2l_ntile1(P)],
                                             const double (&B)[m2l_ntile1(P)], double a)
{
    if constexpr (T < m2l_ntiles(P)) {
        constexpr int ta = M2LTiles<P>::map.ta[T], tb = M2LTiles<P>::map.tb[T];
        C[tb] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, B[ta], C[tb], 0, 0, 0);
    }
}

// Eight tiles: one 16-byte LDS read brings this lane's eight E byte offsets
// (one 16-bit half each, so extracting and adding the row base is a single
// instruction), then eight independent gathers of E feed eight MFMAs.
template <int P, int T0, int... J>
__device__ __forceinline__ void m2l_mfma_batch(double (&C)[m2l_ntile1(P)],
                                               const double (&B)[m2l_ntile1(P)],
                                               const uint16_t *tl, const char *Eb,
                                               std::integer_sequence<int, J...>)
{
    const uint4 q = *reinterpret_cast<const uint4 *>(tl + T0);
    const unsigned w[4] = {q.x, q.y, q.z, q.w};
    double a[8];
    ((a[J] = *reinterpret_cast<const double *>(
          Eb + ((J & 1) ? (w[J >> 1] >> 16) : (w[J >> 1] & 0xffffu)))), ...);
    (m2l_mfma_one<P, T0 + J>(C, B, a[J]), ...);
}

template <int P, int T0>
__device__ __forceinline__ void m2l_mfma_group(double (&C)[m2l_ntile1(P)],
                                               const double (&B)[m2l_ntile1(P)],
                                               const uint16_t *tl, const char *Eb)
{
    if constexpr (T0 < m2l_ntiles(P)) {
        m2l_mfma_batch<P, T0>(C, B, tl, Eb, std::make_integer_sequence<int, 8>{});
        m2l_mfma_group<P, T0 + 8>(C, B, tl, Eb);
    }
}

// Ordering of LDS accesses between the lanes of one wave: a wave's LDS
// operations execute in issue order, so a compiler-level barrier suffices
// (no wait for outstanding global loads, unlike a workgroup barrier).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int m2l_erow(int P) { return (m2l_ncoef(P) + 1 + 31) / 32 * 32 + 8; }

// 1 / ext^n for n = 0 .. M2L_MFMA_MAX_P + 1 (the scales of the M2L's
// dimensionless expansions), by repeated products.
__device__ __forceinline__ void m2l_einv(double ext, double *einv)
{
    double p = 1.0;
    einv[0] = 1.0;
    for (int n = 1; n <= M2L_MFMA_MAX_P + 1; ++n) {
        p *= ext;
        einv[n] = 1.0 / p;
    }
}

// M <- m'' = (-1)^|a| / a! m_a tau^-|a| in place, tau = ext 2^-(l+1) the half
// width of the cell (level l), once the upward pass is done (the multipoles
// have no reader after the M2L).  erec[k].w >> 24 = |k|.
__global__ __launch_bounds__(FMM_BLOCK) void m2l_prescale_kernel(double *__restrict__ M, int64_t n,
                                                                 int K,
                                                                 const double *__restrict__ bscale,
                                                                 const int4 *__restrict__ ijk,
                                                                 const int4 *__restrict__ erec,
                                                                 double ext)
{
    __shared__ double einv[M2L_MFMA_MAX_P + 2];
    if (threadIdx.x == 0) m2l_einv(ext, einv);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i >= n) return;
    const int k = (int)(i % K), dg = erec[k].w >> 24;
    const int l = ijk[i / (3 * K)].w;
    M[i] *= ldexp(bscale[k] * einv[dg], dg * (l + 1));
}

// One entry of the Taylor-coefficient recurrence from its six predecessors
// (written with explicit fused operations, so the table and the kernel's own
// recurrence round alike).
__device__ __forceinline__ double m2l_taylor(double f1, double f2, double r0, double r1, double r2,
                                             double a0, double a1, double a2, double b0, double b1,
                                             double b2)
{
    return fma(f1, fma(r2, a2, fma(r1, a1, r0 * a0)), f2 * ((b0 + b1) + b2));
}

// E_k = k! 2^(dA (|k| + 1)) t_k: the factor from the record word k! | |k| << 24.
__device__ __forceinline__ double m2l_kscale(int w, int dA)
{
    return ldexp((double)(w & 0xffffff), dA * ((w >> 24) + 1));
}

__host__ __device__ constexpr int64_t m2l_tab_row(int dA, int vx, int vy, int vz)
{
    return (((int64_t)dA * M2L_TAB_W + vx + M2L_TAB_V) * M2L_TAB_W + vy + M2L_TAB_V) * M2L_TAB_W +
           vz + M2L_TAB_V;
}

// Geometry of an M2L pair from the cells' (ix, iy, iz, level): separation
// (ext 2^-(lf+1)) r, r integer (exact in FP64), and the pair's table row.
struct M2LPair {
    double r0, r1, r2;
    int dA;
    bool hit;
    int64_t row;
};

__device__ __forceinline__ M2LPair m2l_pair(const int4 a, const int4 s)
{
    const int lf = max(a.w, s.w), da = lf - a.w, ds = lf - s.w;
    const long long vx = ((2ll * a.x + 1) << da) - ((2ll * s.x + 1) << ds);
    const long long vy = ((2ll * a.y + 1) << da) - ((2ll * s.y + 1) << ds);
    const long long vz = ((2ll * a.z + 1) << da) - ((2ll * s.z + 1) << ds);
    M2LPair p;
    p.r0 = (double)vx;
    p.r1 = (double)vy;
    p.r2 = (double)vz;
    p.dA = da;
    p.hit = da <= M2L_TAB_D && llabs(vx) <= M2L_TAB_V && llabs(vy) <= M2L_TAB_V &&
            llabs(vz) <= M2L_TAB_V;
    p.row = p.hit ? m2l_tab_row(da, (int)vx, (int)vy, (int)vz) : 0;
    return p;
}

// Table rows E_k(v) 2^(dA (|k| + 1)), k < K, one thread per row (the zero
// offset row holds zeros), by the recurrence of m2l_mfma_kernel.
template <int P>
__global__ __launch_bounds__(FMM_BLOCK) void m2l_etab_kernel(const int4 *__restrict__ erec,
                                                             double *__restrict__ tab)
{
    constexpr int K = m2l_ncoef(P);
    const int64_t row = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (row >= M2L_TAB_ROWS) return;
    int64_t q = row;
    const int vz = (int)(q % M2L_TAB_W) - M2L_TAB_V;
    q /= M2L_TAB_W;
    const int vy = (int)(q % M2L_TAB_W) - M2L_TAB_V;
    q /= M2L_TAB_W;
    const int vx = (int)(q % M2L_TAB_W) - M2L_TAB_V;
    const int dA = (int)(q / M2L_TAB_W);
    double *t = tab + row * K;
    if (vx == 0 && vy == 0 && vz == 0) {
        for (int k = 0; k < K; ++k) t[k] = 0.0;
        return;
    }
    const double r0 = vx, r1 = vy, r2 = vz;
    const double R2 = r0 * r0 + r1 * r1 + r2 * r2, iR2 = 1.0 / R2;
    t[0] = 1.0 / sqrt(R2);
    auto ld = [&](unsigned b) { return b == 8u * K ? 0.0 : t[b >> 3]; };
    for (int n = 1; n <= P; ++n) {
        const double f1 = -(2.0 * n - 1.0) / n * iR2, f2 = -(n - 1.0) / n * iR2;
        for (int k = m2l_ncoef(n - 1); k < m2l_ncoef(n); ++k) {
            const int4 e = erec[k];
            const unsigned qx = e.x, qy = e.y, qz = e.z;
            t[k] = m2l_taylor(f1, f2, r0, r1, r2, ld(qx & 0xffffu), ld(qx >> 16), ld(qy & 0xffffu),
                              ld(qy >> 16), ld(qz & 0xffffu), ld(qz >> 16));
        }
    }
    for (int k = 0; k < K; ++k) t[k] *= m2l_kscale(erec[k].w, dA);
}

// Degree of entry 4 t + h (h = 0..3) of a column or row tile: base[t] +
// (h >= step[t]) (a tile spans at most one degree boundary).
template <int P>
struct M2LDeg {
    struct Map {
        int base[m2l_ntile1(P)], step[m2l_ntile1(P)];
    };
    static constexpr Map make()
    {
        Map m{};
        for (int t = 0; t < m2l_ntile1(P); ++t) {
            m.base[t] = m2l_degree(4 * t);
            m.step[t] = 4;
            for (int h = 3; h >= 1; --h)
                if (4 * t + h < m2l_ncoef(P) && m2l_degree(4 * t + h) > m.base[t]) m.step[t] = h;
        }
        return m;
    }
    static constexpr Map map = make();
};

template <int P>
__device__ __forceinline__ int m2l_tile_degree(int t, int h)
{
    return M2LDeg<P>::map.base[t] + (h >= M2LDeg<P>::map.step[t] ? 1 : 0);
}

// Persistent: block b serves cells k G + (b mod 8) W + b / 8 (k = 0, 1, ...;
// G = gridDim.x = 8 W), so each XCD (workgroups go round-robin over the 8
// XCDs) sweeps contiguous runs of cells whose sources overlap in its L2, and
// the tables are staged in LDS once per block.  `erec` packs, per k, the
// byte offsets in a row of k - e_i and k - 2 e_i (8 K where absent: the zero
// slot) and k! | |k| << 24: (j1x | j1y << 16, j1z | j2x << 16, j2y | j2z << 16,
// k! | |k| << 24).
//
// Units: with u = ext 2^-(lf+1) (half the finer cell's width) a pair's
// separation is u v, v integer, and E_k(u v) = u^-(|k|+1) E_k(v).  The
// multipoles arrive as m''_a = m'_a tau_S^-|a| (m2l_prescale_kernel) and
// the kernel accumulates lambda_b = sum_a Ehat_{a+b} m''_a 2^((l_A - l_S)|a|)
// with Ehat_k = 2^(dA (|k|+1)) E_k(v) (dA = lf - l_A): l'_b =
// tau_A^-(|b|+1) lambda_b, tau = the half width of each cell.  Ehat is copied
// from the operator table when all four sources of a group have rows there,
// and otherwise built by the recurrence from v; the powers of two are exact.
//
// The four sources' E are rows of m2l_erow(P) doubles: K + 1 (the zero slot at
// K), rounded so that the rows start 16 banks apart (the layout measured
// fastest among the ones tried).  Absent sources and the zero fourth B column
// read the multipole row `zrow` (all zeros).
template <int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(P <= 8 ? 2 : 1)))
void m2l_mfma_kernel(
    const int4 *__restrict__ ijk, int ncell, const int64_t *__restrict__ off,
    const int32_t *__restrict__ src, const int4 *__restrict__ erec,
    const uint16_t *__restrict__ tile_idx, const double *__restrict__ lscale,
    const double *__restrict__ etab, double ext, const double *__restrict__ M, int64_t zrow,
    double *__restrict__ L)
{
    constexpr int K = m2l_ncoef(P), NT = m2l_ntile1(P), N8 = m2l_ntiles8(P), ER = m2l_erow(P);
    constexpr int MAXE = (K + 15) / 16;   // table entries per lane of a source's 16
    __shared__ double E[4 * ER];
    __shared__ __align__(16) uint16_t tab[16 * N8];
    __shared__ int4 rec[K];
    __shared__ double einv[M2L_MFMA_MAX_P + 2];
    __shared__ double tsc[M2L_MFMA_MAX_P + 2];   // tau_A^-n of the current cell
    const int lane = threadIdx.x;
    for (int i = lane; i < 16 * N8; i += 64) tab[i] = tile_idx[i];
    for (int i = lane; i < K; i += 64) rec[i] = erec[i];
    if (lane < 4) E[lane * ER + K] = 0.0;
    if (lane == 0) m2l_einv(ext, einv);
    __syncthreads();
    const int nq = lane & 3, blk = (lane >> 2) & 3, hi = lane >> 4, sub = lane & 15;
    const uint16_t *tl = tab + (nq + 4 * hi) * N8;   // A element (row = lane & 3, k = lane >> 4)
    const char *Eb = reinterpret_cast<const char *>(E + blk * ER);
    double *Ej = E + hi * ER;                   // E of source hi, built by lanes 16 hi ..
    const char *Ejb = reinterpret_cast<const char *>(Ej);
    const int G = gridDim.x, W = G >> 3;
    const int first = (int)(blockIdx.x & 7) * W + (int)(blockIdx.x >> 3);
    for (int c = first; c < ncell; c += G) {
        const int4 A = ijk[c];
        wave_lds_sync();   // the previous cell's reads of tsc are done
        if (lane <= P + 1) tsc[lane] = ldexp(einv[lane], lane * (A.w + 1));
        double Cacc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) Cacc[t] = 0.0;
        const int64_t e0 = off[c], e1 = off[c + 1];
        // Source indices run two groups ahead and cell coordinates one group
        // ahead, so no wait for them lands behind the burst of B loads (loads
        // complete in issue order).  Indices past the list are clamped to its
        // last entry.
        const int64_t elast = e1 - 1;
        int ib = 0, ib1 = 0, ie1 = 0;
        int4 S = make_int4(0, 0, 0, 0);
        if (e0 < e1) {
            ib = src[min(e0 + blk, elast)];
            ib1 = src[min(e0 + 4 + blk, elast)];
            ie1 = src[min(e0 + 4 + hi, elast)];
            S = ijk[src[min(e0 + hi, elast)]];
        }
        for (int64_t eb = e0; eb < e1; eb += 4) {
            const int nb = (int)min<int64_t>(4, e1 - eb);
            const int4 Sn = ijk[ie1];   // next group's coordinates of source hi
            const int ie2 = src[min(eb + 8 + hi, elast)], ib2 = src[min(eb + 8 + blk, elast)];
            // geometry of source hi (absent sources: v = (2, 0, 0), finite E, zero B)
            M2LPair g = m2l_pair(A, S);
            if (hi >= nb) {
                g.r0 = 2.0;
                g.r1 = 0.0;
                g.r2 = 0.0;
                g.dA = 0;
                g.hit = true;
                g.row = m2l_tab_row(0, 2, 0, 0);
            }
            const bool tabled = __all(g.hit);
            double ev[MAXE];
            if (tabled) {
                const double *er = etab + g.row * K;
#pragma unroll
                for (int it = 0; it < MAXE; ++it) ev[it] = er[min(sub + 16 * it, K - 1)];
            }
            const int lsb = ijk[ib].w;   // level of source blk
            // B operand: m''_blk[4 ta + k][n], k = lane >> 4, n = lane & 3.
            // Every lane loads from a valid address and multiplies by a 0/1
            // mask, so the loads issue back to back instead of one guarded
            // load at a time (the multipoles are finite).
            double Bv[NT];
            const double *Mb =
                M + ((blk < nb && nq < 3) ? (size_t)ib * 3 + nq : (size_t)zrow * 3) * K;
#pragma unroll
            for (int ta = 0; ta < NT - 1; ++ta) Bv[ta] = Mb[4 * ta + hi];
            Bv[NT - 1] = Mb[min(4 * (NT - 1) + hi, K - 1)] * (4 * (NT - 1) + hi < K ? 1.0 : 0.0);
            S = Sn;
            ie1 = ie2;
            ib = ib1;
            ib1 = ib2;
            wave_lds_sync();   // the previous group's reads of E are done
            if (tabled) {
#pragma unroll
                for (int it = 0; it < MAXE; ++it) {
                    const int k = sub + 16 * it;
                    if (k < K) Ej[k] = ev[it];
                }
            } else {
                const double r0 = g.r0, r1 = g.r1, r2 = g.r2;
                const double R2 = r0 * r0 + r1 * r1 + r2 * r2, iR2 = 1.0 / R2;
                if (sub == 0) Ej[0] = 1.0 / sqrt(R2);
                auto ld = [&](unsigned b) { return *reinterpret_cast<const double *>(Ejb + b); };
                constexpr int MAXIT = ((P + 1) * (P + 2) / 2 + 15) / 16;
#pragma unroll
                for (int n = 1; n <= P; ++n) {
                    wave_lds_sync();
                    const double f1 = -(2.0 * n - 1.0) / n * iR2, f2 = -(n - 1.0) / n * iR2;
#pragma unroll
                    for (int it = 0; it < MAXIT; ++it) {
                        const int k = m2l_ncoef(n - 1) + sub + 16 * it;
                        if (k < m2l_ncoef(n)) {
                            const int4 q = rec[k];
                            const unsigned qx = q.x, qy = q.y, qz = q.z;
                            Ej[k] = m2l_taylor(f1, f2, r0, r1, r2, ld(qx & 0xffffu), ld(qx >> 16),
                                               ld(qy & 0xffffu), ld(qy >> 16), ld(qz & 0xffffu),
                                               ld(qz >> 16));
                        }
                    }
                }
                wave_lds_sync();
                for (int k = sub; k < K; k += 16) Ej[k] *= m2l_kscale(rec[k].w, g.dA);
            }
            // level gap of source blk: m''_a 2^((l_A - l_S) |a|), a = 4 ta + hi
            const int eB = A.w - lsb;
            if (__any(eB != 0)) {
#pragma unroll
                for (int ta = 0; ta < NT; ++ta) Bv[ta] = ldexp(Bv[ta], eB * m2l_tile_degree<P>(ta, hi));
            }
            wave_lds_sync();
            m2l_mfma_group<P, 0>(Cacc, Bv, tl, Eb);
        }
        // sum the four blocks (lanes 4 and 8 apart), scale by tau_A^-(|b|+1) / b!,
        // write l
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            double v = Cacc[t];
            v += __shfl_xor(v, 4);
            v += __shfl_xor(v, 8);
            const int b = 4 * t + hi, dg = m2l_tile_degree<P>(t, hi);
            const double ls = lscale[min(b, K - 1)] * tsc[dg + 1];
            if (blk == 0 && nq < 3 && b < K) L[((size_t)c * 3 + nq) * K + b] = v * ls;
        }
    }
}

// L2L for the children listed in `kids` (one level): l_g(child) +=
// sum_{b >= g} C(b,g) l_b(parent) e^{b-g}, e = c_child - c_parent.
__global__ __launch_bounds__(FMM_BLOCK) void l2l_kernel(const int32_t *__restrict__ kids,
                                                        const int32_t *__restrict__ parent,
                                                        const FmmCellDev *__restrict__ cells,
                                                        int K, const int *__restrict__ mi,
                                                        const int *__restrict__ toff,
                                                        const int2 *__restrict__ tidx,
                                                        const double *__restrict__ tcoef,
                                                        double *__restrict__ L)
{
    extern __shared__ double sh[];
    double *mono = sh;
    double *lp = sh + K;
    const int d = kids[blockIdx.x];
    const int c = parent[d];
    const FmmCellDev D = cells[d], C = cells[c];
    const double e[3] = {D.cx - C.cx, D.cy - C.cy, D.cz - C.cz};
    lds_monomials(e, K, mi, mono);
    for (int k = threadIdx.x; k < 3 * K; k += blockDim.x) lp[k] = L[(size_t)c * 3 * K + k];
    __syncthreads();
    for (int r = 0; r < 2; r++) {
        const int k = threadIdx.x + r * FMM_BLOCK;
        if (k >= K) continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int t = toff[k]; t < toff[k + 1]; t++) {
            const int2 ix = tidx[t];   // (b, b-g)
            const double w = tcoef[t] * mono[ix.y];
            s0 += lp[ix.x] * w;
            s1 += lp[K + ix.x] * w;
            s2 += lp[2 * K + ix.x] * w;
        }
        L[((size_t)d * 3 + 0) * K + k] += s0;
        L[((size_t)d * 3 + 1) * K + k] += s1;
        L[((size_t)d * 3 + 2) * K + k] += s2;
    }
}

// L2P: one block per leaf, one particle per thread; far-field U and J.
// Writes U, J at the particle's original index (overwrite, or add when
// reset == 0).
__global__ __launch_bounds__(P2P_BLOCK) void l2p_kernel(const int32_t *__restrict__ leaves,
                                                        const FmmCellDev *__restrict__ cells,
                                                        const double4 *__restrict__ rec,
                                                        const int32_t *__restrict__ perm, int K,
                                                        int P, const int *__restrict__ mi,
                                                        const int4 *__restrict__ d1idx,
                                                        const int4 *__restrict__ d2idx_a,
                                                        const int4 *__restrict__ d2idx_b,
                                                        const double *__restrict__ L,
                                                        int64_t cap, double *__restrict__ U,
                                                        double *__restrict__ J, int reset)
{
    extern __shared__ double sh[];
    double *lcell = sh;                         // 3K
    double *pw = sh + 3 * K;                    // P2P_BLOCK x 3 x (P+1)
    const int c = leaves[blockIdx.x];
    const FmmCellDev C = cells[c];
    for (int k = threadIdx.x; k < 3 * K; k += blockDim.x) lcell[k] = L[(size_t)c * 3 * K + k];
    __syncthreads();
    const int stride = 3 * (P + 1);
    double *mp = pw + threadIdx.x * stride;
    for (int base = C.begin; base < C.end; base += P2P_BLOCK) {
        const int j = base + threadIdx.x;
        if (j >= C.end) continue;
        const double4 a = rec[2 * j];
        const double t[3] = {a.x - C.cx, a.y - C.cy, a.z - C.cz};
        for (int i = 0; i < 3; i++) {
            mp[i * (P + 1)] = 1.0;
            for (int e = 1; e <= P; e++) mp[i * (P + 1) + e] = mp[i * (P + 1) + e - 1] * t[i];
        }
        double d1[3][3], d2[3][6];
        for (int q = 0; q < 3; q++) {
            for (int i = 0; i < 3; i++) d1[q][i] = 0.0;
            for (int i = 0; i < 6; i++) d2[q][i] = 0.0;
        }
        for (int kb = 0; kb < K; kb++) {
            const int e0 = mi[3 * kb], e1 = mi[3 * kb + 1], e2 = mi[3 * kb + 2];
            const double mono = mp[e0] * mp[(P + 1) + e1] * mp[2 * (P + 1) + e2];
            const int4 k1 = d1idx[kb];    // idx(b + e_j), j = 0..2
            const int4 ka = d2idx_a[kb];  // idx(b + e_j + e_l) for (0,0) (0,1) (0,2)
            const int4 kc = d2idx_b[kb];  // for (1,1) (1,2) (2,2)
            const int bexp[3] = {e0, e1, e2};
            const int k1a[3] = {k1.x, k1.y, k1.z};
            for (int jj = 0; jj < 3; jj++) {
                if (k1a[jj] < 0) continue;
                const double f1 = (double)(bexp[jj] + 1) * mono;
                for (int q = 0; q < 3; q++) d1[q][jj] += lcell[q * K + k1a[jj]] * f1;
            }
            // second derivatives: (j,l) pairs in order (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
            const int k2a[6] = {ka.x, ka.y, ka.z, kc.x, kc.y, kc.z};
            const int pj[6] = {0, 0, 0, 1, 1, 2}, pl[6] = {0, 1, 2, 1, 2, 2};
            for (int m = 0; m < 6; m++) {
                if (k2a[m] < 0) continue;
                const int jj = pj[m], ll = pl[m];
                const double f2 = (double)(bexp[jj] + 1) * (double)(bexp[ll] + (jj == ll) + 1) * mono;
                for (int q = 0; q < 3; q++) d2[q][m] += lcell[q * K + k2a[m]] * f2;
            }
        }
        // symmetric second derivatives h[q][j][l]
        auto h = [&](int q, int jj, int ll) -> double {
            if (jj > ll) { const int t2 = jj; jj = ll; ll = t2; }
            const int m = jj == 0 ? ll : (jj == 1 ? 2 + ll : 5);
            return d2[q][m];
        };
        double u[3], Jm[9];
        u[0] = d1[2][1] - d1[1][2];
        u[1] = d1[0][2] - d1[2][0];
        u[2] = d1[1][0] - d1[0][1];
        for (int l = 0; l < 3; l++) {
            Jm[0 + 3 * l] = h(2, 1, l) - h(1, 2, l);
            Jm[1 + 3 * l] = h(0, 2, l) - h(2, 0, l);
            Jm[2 + 3 * l] = h(1, 0, l) - h(0, 1, l);
        }
        const int64_t i = perm[j];
        for (int k = 0; k < 3; k++) {
            const double v = INV4PI * u[k];
            U[k * cap + i] = reset ? v : U[k * cap + i] + v;
        }
        for (int k = 0; k < 9; k++) {
            const double v = INV4PI * Jm[k];
            J[k * cap + i] = reset ? v : J[k * cap + i] + v;
        }
    }
}

// Stores of the near-field kernels, shared with the reduction of partial sums:
// FP64 P2P sums d = (U, J) added to U, J.
__device__ __forceinline__ void p2p_f64_finish(const double d[12], int64_t i, int64_t cap,
                                               double *__restrict__ U, double *__restrict__ J)
{
    for (int k = 0; k < 3; ++k) U[k * cap + i] += d[k];
    for (int k = 0; k < 9; ++k) J[k * cap + i] += d[3 + k];
}

// Mixed P2P sums d = (U, W, J0..7) before -1/(4 pi): U and J added to U, J
// (W the Kronecker part of J, J8 from the zero trace).
__device__ __forceinline__ void p2p_mixed_finish(const double d[14], int64_t i, int64_t cap,
                                                 double *__restrict__ U, double *__restrict__ J)
{
    const double k = -INV4PI;
    U[i] += k * d[0]; U[cap + i] += k * d[1]; U[2 * cap + i] += k * d[2];
    J[i] += k * d[6];
    J[cap + i] += k * (d[7] - d[5]);
    J[2 * cap + i] += k * (d[8] + d[4]);
    J[3 * cap + i] += k * (d[9] + d[5]);
    J[4 * cap + i] += k * d[10];
    J[5 * cap + i] += k * (d[11] - d[3]);
    J[6 * cap + i] += k * (d[12] - d[4]);
    J[7 * cap + i] += k * (d[13] + d[3]);
    J[8 * cap + i] += -k * (d[6] + d[10]);
}

// P2P near field, FP64: one block (one wave) per work item (a target leaf and
// a run of its list), one target per lane, source leaves staged through LDS.
// Adds to U, J (or writes partial sums).  DS: the core-size derivatives D[U],
// D[J] instead (g, g' -> h, h'), added to U, J as given.
template <int KK, bool DS = false>
__global__ __launch_bounds__(P2P_BLOCK) void p2p_f64_kernel(const P2PChunk *__restrict__ chunks,
                                                            const FmmCellDev *__restrict__ cells,
                                                            const int32_t *__restrict__ src,
                                                            const double4 *__restrict__ rec,
                                                            const int32_t *__restrict__ perm,
                                                            int64_t cap, double *__restrict__ U,
                                                            double *__restrict__ J,
                                                            double *__restrict__ part,
                                                            int64_t pstride)
{
    __shared__ double4 tile[2 * P2P_BLOCK];
    const P2PChunk ch = chunks[blockIdx.x];
    const int c = ch.c;
    const FmmCellDev A = cells[c];
    for (int tb = A.begin; tb < A.end; tb += P2P_BLOCK) {
        const int it = tb + threadIdx.x;
        const bool active = it < A.end;
        double xi = 0, yi = 0, zi = 0;
        if (active) {
            const double4 a = rec[2 * it];
            xi = a.x; yi = a.y; zi = a.z;
        }
        double u0 = 0, u1 = 0, u2 = 0, j0 = 0, j1 = 0, j2 = 0, j3 = 0, j4 = 0, j5 = 0, j6 = 0,
               j7 = 0, j8 = 0;
        for (int64_t e = ch.e0; e < ch.e1; e++) {
            const FmmCellDev B = cells[src[e]];
            for (int sb = B.begin; sb < B.end; sb += P2P_BLOCK) {
                __syncthreads();
                const int js = sb + threadIdx.x;
                if (js < B.end) {
                    tile[2 * threadIdx.x] = rec[2 * js];
                    tile[2 * threadIdx.x + 1] = rec[2 * js + 1];
                }
                __syncthreads();
                const int cnt = min(P2P_BLOCK, B.end - sb);
                for (int jj = 0; jj < cnt; ++jj) {
                    const double4 a = tile[2 * jj], b = tile[2 * jj + 1];
                    const double dX1 = xi - a.x, dX2 = yi - a.y, dX3 = zi - a.z;
                    const double r2 = dX1 * dX1 + dX2 * dX2 + dX3 * dX3;
                    if (r2 > 0.0) {
                        const double r = sqrt(r2);
                        double g, dg;
                        if constexpr (DS)
                            h_dh_f64<KK>(r * b.w, g, dg);
                        else
                            g_dg_f64<KK>(r * b.w, g, dg);
                        const double rinv = 1.0 / r;
                        const double r3i = rinv * rinv * rinv;
                        const double G1 = a.w, G2 = b.x, G3 = b.y;
                        const double c1 = -INV4PI * r3i * (dX2 * G3 - dX3 * G2);
                        const double c2 = -INV4PI * r3i * (dX3 * G1 - dX1 * G3);
                        const double c3 = -INV4PI * r3i * (dX1 * G2 - dX2 * G1);
                        u0 += g * c1; u1 += g * c2; u2 += g * c3;
                        const double aux = dg * b.w * rinv - 3.0 * g * rinv * rinv;
                        j0 += aux * c1 * dX1; j1 += aux * c2 * dX1; j2 += aux * c3 * dX1;
                        j3 += aux * c1 * dX2; j4 += aux * c2 * dX2; j5 += aux * c3 * dX2;
                        j6 += aux * c1 * dX3; j7 += aux * c2 * dX3; j8 += aux * c3 * dX3;
                        const double a2 = -INV4PI * g * r3i;
                        j1 -= a2 * G3; j2 += a2 * G2; j3 += a2 * G3;
                        j5 -= a2 * G1; j6 -= a2 * G2; j7 += a2 * G1;
                    }
                }
            }
        }
        if (active) {
            const double d[12] = {u0, u1, u2, j0, j1, j2, j3, j4, j5, j6, j7, j8};
            if (ch.part < 0)
                p2p_f64_finish(d, perm[it], cap, U, J);
            else
                for (int k = 0; k < 12; ++k) part[k * pstride + ch.part + (it - A.begin)] = d[k];
        }
    }
}

// P2P near field, mixed: FP32 pair math on coordinates relative to the
// target leaf's centre, FP32 partials per source tile folded into FP64.
//
// One wave per target leaf.  The particles of the leaf's source leaves are
// staged through LDS as one concatenated stream, 64 records per tile whatever
// the leaf boundaries (a prefix of the source-leaf sizes, 64 leaves at a time,
// maps a stream position to its particle).  A leaf of nt <= 32 targets is
// replicated R = 2, 4 or 8 times across the wave: lane = s*nt + t works on
// target t with the sources s, s+R, s+2R, ... of every tile, and the R partial
// sums of a target are added in slice order at the end (deterministic).
constexpr int P2P_SLOTS = 64;

// Source-leaf stream of one slot (up to P2P_SLOTS leaves from the list at eb):
// lbeg = each leaf's first particle, lpre = inclusive prefix of the leaf sizes
// (a wave scan; the block is one wave).  Returns the slot's particle count.
__device__ __forceinline__ int stream_setup(const FmmCellDev *__restrict__ cells,
                                            const int32_t *__restrict__ src, int64_t eb, int nl,
                                            int lane, int32_t *lbeg, int32_t *lpre)
{
    int v = 0;
    if (lane < nl) {
        const FmmCellDev B = cells[src[eb + lane]];
        lbeg[lane] = B.begin;
        v = B.end - B.begin;
    }
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    lpre[lane + 1] = v;
    if (lane == 0) lpre[0] = 0;
    __syncthreads();
    return lpre[nl];
}

// Particle at stream position q < total of the slot.
__device__ __forceinline__ int stream_particle(const int32_t *lbeg, const int32_t *lpre, int nl,
                                               int q)
{
    int lo = 0, hi = nl;   // lpre[lo] <= q < lpre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lpre[mid] <= q) lo = mid; else hi = mid;
    }
    return lbeg[lo] + (q - lpre[lo]);
}

template <int KK, int R, bool DS>
__device__ __forceinline__ void p2p_mixed_tile(const float4 *tpos, const float4 *tgam, int cnt,
                                               int s, float xi, float yi, float zi, MixedAcc &acc)
{
    if (cnt == P2P_BLOCK) {
#pragma unroll 4
        for (int i = 0; i < P2P_BLOCK / R; ++i) {
            const int j = i * R + s;
            if constexpr (DS)
                mixed_pair_dsig<KK>(tpos[j], tgam[j], xi, yi, zi, acc);
            else
                mixed_pair<KK>(tpos[j], tgam[j], xi, yi, zi, acc);
        }
    } else {
        for (int j = s; j < cnt; j += R) {
            if constexpr (DS)
                mixed_pair_dsig<KK>(tpos[j], tgam[j], xi, yi, zi, acc);
            else
                mixed_pair<KK>(tpos[j], tgam[j], xi, yi, zi, acc);
        }
    }
}

// DS: the core-size derivatives D[U], D[J] in the same sums (mixed_pair_dsig).
template <int KK, bool DS = false>
__global__ __launch_bounds__(P2P_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void p2p_mixed_kernel(const P2PChunk *__restrict__ chunks,
                                                              const FmmCellDev *__restrict__ cells,
                                                              const int32_t *__restrict__ src,
                                                              const double4 *__restrict__ rec,
                                                              const int32_t *__restrict__ perm,
                                                              int64_t cap, double *__restrict__ U,
                                                              double *__restrict__ J,
                                                              double *__restrict__ part,
                                                              int64_t pstride)
{
    __shared__ float4 tpos[P2P_BLOCK];
    __shared__ float4 tgam[P2P_BLOCK];
    __shared__ int32_t lbeg[P2P_SLOTS], lpre[P2P_SLOTS + 1];
    __shared__ double red[14][P2P_BLOCK];
    const int lane = threadIdx.x;
    const P2PChunk ch = chunks[blockIdx.x];
    const int c = ch.c;
    const FmmCellDev A = cells[c];
    const int64_t e0 = ch.e0, e1 = ch.e1;
    for (int tb = A.begin; tb < A.end; tb += P2P_BLOCK) {
        const int nt = min(P2P_BLOCK, A.end - tb);
        const int R = nt <= 8 ? 8 : nt <= 16 ? 4 : nt <= 32 ? 2 : 1;
        const int sl = lane / nt;                  // slice of this lane
        const int t = lane - sl * nt;              // target of this lane
        const bool active = sl < R;
        float xi = 0.f, yi = 0.f, zi = 0.f;
        {
            const double4 a = rec[2 * (tb + t)];
            xi = (float)(a.x - A.cx); yi = (float)(a.y - A.cy); zi = (float)(a.z - A.cz);
        }
        const int s = active ? sl : 0;
        double dU0 = 0, dU1 = 0, dU2 = 0, dW0 = 0, dW1 = 0, dW2 = 0;
        double dJ0 = 0, dJ1 = 0, dJ2 = 0, dJ3 = 0, dJ4 = 0, dJ5 = 0, dJ6 = 0, dJ7 = 0;
        for (int64_t eb = e0; eb < e1; eb += P2P_SLOTS) {
            const int nl = (int)min<int64_t>(P2P_SLOTS, e1 - eb);
            __syncthreads();
            const int total = stream_setup(cells, src, eb, nl, lane, lbeg, lpre);
            // the records of the next tile load while this one computes
            double4 ra = make_double4(0.0, 0.0, 0.0, 0.0), rb = ra;
            if (lane < total) {
                const int js = stream_particle(lbeg, lpre, nl, lane);
                ra = rec[2 * js];
                rb = rec[2 * js + 1];
            }
            for (int base = 0; base < total; base += P2P_BLOCK) {
                __syncthreads();
                if (base + lane < total) {
                    float p0 = 0.f, p1 = 0.f;
                    if constexpr (KK == GAUSSIANERF || KK == GAUSSIAN) {
                        p0 = (float)rb.w;
                        p1 = (float)(rb.w * rb.w);
                    } else if constexpr (KK == WINCKELMANS) {
                        p0 = (float)(rb.z * rb.z);
                    }
                    tpos[lane] = make_float4((float)(ra.x - A.cx), (float)(ra.y - A.cy),
                                             (float)(ra.z - A.cz), p0);
                    tgam[lane] = make_float4((float)ra.w, (float)rb.x, (float)rb.y, p1);
                }
                __syncthreads();
                const int qn = base + P2P_BLOCK + lane;
                if (qn < total) {
                    const int js = stream_particle(lbeg, lpre, nl, qn);
                    ra = rec[2 * js];
                    rb = rec[2 * js + 1];
                }
                const int cnt = min(P2P_BLOCK, total - base);
                MixedAcc acc;
                switch (R) {
                case 8: p2p_mixed_tile<KK, 8, DS>(tpos, tgam, cnt, s, xi, yi, zi, acc); break;
                case 4: p2p_mixed_tile<KK, 4, DS>(tpos, tgam, cnt, s, xi, yi, zi, acc); break;
                case 2: p2p_mixed_tile<KK, 2, DS>(tpos, tgam, cnt, s, xi, yi, zi, acc); break;
                default: p2p_mixed_tile<KK, 1, DS>(tpos, tgam, cnt, s, xi, yi, zi, acc); break;
                }
                dU0 += acc.U0; dU1 += acc.U1; dU2 += acc.U2;
                dW0 += acc.W0; dW1 += acc.W1; dW2 += acc.W2;
                dJ0 += acc.J0; dJ1 += acc.J1; dJ2 += acc.J2; dJ3 += acc.J3;
                dJ4 += acc.J4; dJ5 += acc.J5; dJ6 += acc.J6; dJ7 += acc.J7;
            }
        }
        if (R > 1) {   // add the slices of each target in slice order
            const double v[14] = {dU0, dU1, dU2, dW0, dW1, dW2, dJ0, dJ1, dJ2, dJ3, dJ4, dJ5, dJ6, dJ7};
            __syncthreads();
            for (int k = 0; k < 14; ++k) red[k][lane] = v[k];
            __syncthreads();
            if (lane < nt) {
                double w[14];
                for (int k = 0; k < 14; ++k) {
                    double acc2 = red[k][lane];
                    for (int q = 1; q < R; ++q) acc2 += red[k][lane + q * nt];
                    w[k] = acc2;
                }
                dU0 = w[0]; dU1 = w[1]; dU2 = w[2]; dW0 = w[3]; dW1 = w[4]; dW2 = w[5];
                dJ0 = w[6]; dJ1 = w[7]; dJ2 = w[8]; dJ3 = w[9]; dJ4 = w[10]; dJ5 = w[11];
                dJ6 = w[12]; dJ7 = w[13];
            }
        }
        if (lane < nt) {
            const double d[14] = {dU0, dU1, dU2, dW0, dW1, dW2, dJ0, dJ1, dJ2, dJ3, dJ4, dJ5, dJ6, dJ7};
            if (ch.part < 0)
                p2p_mixed_finish(d, perm[tb + lane], cap, U, J);
            else
                for (int k = 0; k < 14; ++k)
                    part[k * pstride + ch.part + (tb + lane - A.begin)] = d[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Host: tree build and traversal (same rules as the oracle's CPU FMM)
// ---------------------------------------------------------------------------

static int64_t lower_bound_key(const uint64_t *key, int64_t b, int64_t e, uint64_t v)
{
    while (b < e) {
        int64_t m = b + (e - b) / 2;
        if (key[m] < v) b = m + 1; else e = m;
    }
    return b;
}

// fn(begin, end) over [0, n) on up to 16 host threads (serially when n is small).
template <class F>
static void host_parallel_for(int64_t n, F fn)
{
    const int64_t grain = 4096;
    int nt = (int)std::min<int64_t>((n + grain - 1) / grain,
                                    std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    if (nt <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 1; t < nt; t++) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) th.emplace_back(fn, b, e);
    }
    fn((int64_t)0, std::min(n, chunk));
    for (auto &x : th) x.join();
}

// Octree from the sorted keys, level by level: the cells of a level split in
// parallel (each non-empty octant of a cell with more than ncrit particles is
// a child, its range found by binary search in the parent's key range), and
// the children are appended in the order of their parents and octants -- the
// breadth-first order of a sequential queue, so the tree does not depend on
// the thread count.
void fmm_build_tree(FmmPlan &pl, const uint64_t *key, const double *ssig, int64_t n, int ncrit)
{
    pl.cells.clear();
    FmmCellHost root{};
    root.begin = 0; root.end = n; root.level = 0; root.child0 = -1; root.nchild = 0;
    root.prefix = 0;
    pl.cells.push_back(root);
    struct Kids {
        int n;
        int64_t b[8], e[8];
    };
    std::vector<Kids> kids;
    for (int64_t lb = 0, le = 1; lb < le; lb = le, le = (int64_t)pl.cells.size()) {
        kids.assign(le - lb, Kids{});
        host_parallel_for(le - lb, [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; i++) {
                const FmmCellHost &C = pl.cells[lb + i];
                Kids &K = kids[i];
                K.n = 0;
                if (C.end - C.begin <= ncrit || C.level >= KEYBITS) continue;
                const int shift = 3 * (KEYBITS - C.level - 1);
                for (uint64_t o = 0; o < 8; o++) {
                    const uint64_t cp = (C.prefix << 3) | o;
                    const int64_t b = lower_bound_key(key, C.begin, C.end, cp << shift);
                    const int64_t e = lower_bound_key(key, C.begin, C.end, (cp + 1) << shift);
                    if (e == b) continue;
                    K.b[o] = b;
                    K.e[o] = e;
                    K.n |= 1 << o;
                }
            }
        });
        for (int64_t i = 0; i < le - lb; i++) {
            if (!kids[i].n) continue;
            const int64_t first = (int64_t)pl.cells.size();
            const FmmCellHost P = pl.cells[lb + i];
            int nch = 0;
            for (int o = 0; o < 8; o++) {
                if (!(kids[i].n >> o & 1)) continue;
                FmmCellHost D{};
                D.begin = kids[i].b[o]; D.end = kids[i].e[o]; D.level = P.level + 1;
                D.child0 = -1; D.nchild = 0;
                D.prefix = (P.prefix << 3) | (uint64_t)o;
                pl.cells.push_back(D);
                nch++;
            }
            pl.cells[lb + i].child0 = first;
            pl.cells[lb + i].nchild = nch;
        }
    }
    // geometry
    host_parallel_for((int64_t)pl.cells.size(), [&](int64_t i0, int64_t i1) {
        for (int64_t ci = i0; ci < i1; ci++) {
            FmmCellHost &C = pl.cells[ci];
            const uint64_t pre = C.prefix;
            uint64_t ix = 0, iy = 0, iz = 0;
            for (int l = 0; l < C.level; l++) {
                const uint64_t o = (pre >> (3 * (C.level - 1 - l))) & 7u;
                ix = (ix << 1) | (o & 1u);
                iy = (iy << 1) | ((o >> 1) & 1u);
                iz = (iz << 1) | ((o >> 2) & 1u);
            }
            const double h = pl.ext / (double)((uint64_t)1 << C.level);
            C.c[0] = pl.lo[0] + ((double)ix + 0.5) * h;
            C.c[1] = pl.lo[1] + ((double)iy + 0.5) * h;
            C.c[2] = pl.lo[2] + ((double)iz + 0.5) * h;
            C.R = 0.8660254037844386 * h;
        }
    });
    // sigma max: leaves in parallel, then parents level by level from the
    // deepest (cells of a level are contiguous in breadth-first order)
    const int64_t nc = (int64_t)pl.cells.size();
    host_parallel_for(nc, [&](int64_t i0, int64_t i1) {
        for (int64_t ci = i0; ci < i1; ci++) {
            FmmCellHost &C = pl.cells[ci];
            if (C.nchild) continue;
            double s = 0.0;
            for (int64_t j = C.begin; j < C.end; j++) s = fmax(s, ssig[j]);
            C.smax = s;
        }
    });
    for (int64_t ci = nc - 1; ci >= 0; ci--) {
        FmmCellHost &C = pl.cells[ci];
        if (C.nchild == 0) continue;
        double s = 0.0;
        for (int k = 0; k < C.nchild; k++) s = fmax(s, pl.cells[C.child0 + k].smax);
        C.smax = s;
    }
}

static bool mac(const FmmCellHost &A, const FmmCellHost &B, double theta, double sigcut)
{
    const double dx = A.c[0] - B.c[0], dy = A.c[1] - B.c[1], dz = A.c[2] - B.c[2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    return (A.R + B.R) < theta * d && (d - A.R - B.R) >= sigcut * B.smax;
}

void fmm_traverse(FmmPlan &pl, double theta, double sigcut)
{
    std::vector<std::pair<int32_t, int32_t>> m2l, p2p, st;
    st.reserve(1 << 16);
    if (!pl.cells.empty()) st.push_back({0, 0});
    const auto &cells = pl.cells;
    while (!st.empty()) {
        const auto pr = st.back();
        st.pop_back();
        const int32_t a = pr.first, b = pr.second;
        const FmmCellHost &A = cells[a], &B = cells[b];
        if (a == b) {
            if (A.nchild == 0) {
                p2p.push_back({a, b});
            } else {
                for (int i = A.nchild - 1; i >= 0; i--)
                    for (int j = A.nchild - 1; j >= 0; j--)
                        st.push_back({(int32_t)(A.child0 + i), (int32_t)(A.child0 + j)});
            }
        } else if (mac(A, B, theta, sigcut)) {
            m2l.push_back({a, b});
        } else if (A.nchild == 0 && B.nchild == 0) {
            p2p.push_back({a, b});
        } else if (B.nchild == 0 || (A.nchild != 0 && A.R >= B.R)) {
            for (int i = A.nchild - 1; i >= 0; i--) st.push_back({(int32_t)(A.child0 + i), b});
        } else {
            for (int j = B.nchild - 1; j >= 0; j--) st.push_back({a, (int32_t)(B.child0 + j)});
        }
    }
    std::sort(m2l.begin(), m2l.end());
    std::sort(p2p.begin(), p2p.end());
    const int64_t nc = (int64_t)cells.size();
    auto csr = [nc](const std::vector<std::pair<int32_t, int32_t>> &v, std::vector<int64_t> &off,
                    std::vector<int32_t> &src) {
        off.assign(nc + 1, 0);
        src.resize(v.size());
        for (size_t i = 0; i < v.size(); i++) {
            off[v[i].first + 1]++;
            src[i] = v[i].second;
        }
        for (int64_t c = 0; c < nc; c++) off[c + 1] += off[c];
    };
    csr(m2l, pl.m2l_off, pl.m2l_src);
    csr(p2p, pl.p2p_off, pl.p2p_src);
}

// Operator tables for expansion order P (host build, uploaded once per P).
void fmm_build_tables(FmmTables &T, int P)
{
    MIndex M;
    M.build(P);
    const int K = M.K;
    T.P = P;
    T.K = K;
    T.mi = M.mi;
    // derivative recurrence: idx(k - e_i), idx(k - 2 e_i)
    T.rec1.assign(4 * K, -1);
    T.rec2.assign(4 * K, -1);
    for (int k = 0; k < K; k++) {
        const int *a = &M.mi[3 * k];
        for (int i = 0; i < 3; i++) {
            int d[3] = {a[0], a[1], a[2]};
            d[i] -= 1;
            T.rec1[4 * k + i] = M.idx(d[0], d[1], d[2]);
            d[i] -= 1;
            T.rec2[4 * k + i] = M.idx(d[0], d[1], d[2]);
        }
    }
    // M2M: for each a, (b, a-b, C(a,b))
    T.m2m_off.assign(K + 1, 0);
    T.m2m_idx.clear();
    T.m2m_coef.clear();
    for (int k = 0; k < K; k++) {
        const int *a = &M.mi[3 * k];
        for (int b0 = 0; b0 <= a[0]; b0++)
            for (int b1 = 0; b1 <= a[1]; b1++)
                for (int b2 = 0; b2 <= a[2]; b2++) {
                    T.m2m_idx.push_back(M.idx(b0, b1, b2));
                    T.m2m_idx.push_back(M.idx(a[0] - b0, a[1] - b1, a[2] - b2));
                    T.m2m_coef.push_back(binom(a[0], b0) * binom(a[1], b1) * binom(a[2], b2));
                }
        T.m2m_off[k + 1] = (int)T.m2m_coef.size();
    }
    // L2L: for each g, (b, b-g, C(b,g)) over b >= g in index order
    T.l2l_off.assign(K + 1, 0);
    T.l2l_idx.clear();
    T.l2l_coef.clear();
    for (int kg = 0; kg < K; kg++) {
        const int *g = &M.mi[3 * kg];
        for (int kb = 0; kb < K; kb++) {
            const int *b = &M.mi[3 * kb];
            if (b[0] < g[0] || b[1] < g[1] || b[2] < g[2]) continue;
            T.l2l_idx.push_back(kb);
            T.l2l_idx.push_back(M.idx(b[0] - g[0], b[1] - g[1], b[2] - g[2]));
            T.l2l_coef.push_back(binom(b[0], g[0]) * binom(b[1], g[1]) * binom(b[2], g[2]));
        }
        T.l2l_off[kg + 1] = (int)T.l2l_coef.size();
    }
    // M2L: for each b, (a, a+b, (-1)^|a| C(a+b, a)), |a| + |b| <= P, a in index order
    T.m2l_off.assign(K + 1, 0);
    T.m2l_idx.clear();
    T.m2l_coef.clear();
    for (int kb = 0; kb < K; kb++) {
        const int *b = &M.mi[3 * kb];
        for (int ka = 0; ka < K; ka++) {
            const int *a = &M.mi[3 * ka];
            const int na = a[0] + a[1] + a[2];
            if (na + b[0] + b[1] + b[2] > P) continue;
            T.m2l_idx.push_back(ka);
            T.m2l_idx.push_back(M.idx(a[0] + b[0], a[1] + b[1], a[2] + b[2]));
            T.m2l_coef.push_back((na % 2 ? -1.0 : 1.0) * binom(a[0] + b[0], a[0]) *
                                 binom(a[1] + b[1], a[1]) * binom(a[2] + b[2], a[2]));
        }
        T.m2l_off[kb + 1] = (int)T.m2l_coef.size();
    }
    // M2L chunks: the length that minimises the per-thread work (rounds x length)
    {
        auto count = [&](int len) {
            int n = 0;
            for (int kb = 0; kb < K; kb++) n += (T.m2l_off[kb + 1] - T.m2l_off[kb] + len - 1) / len;
            return n;
        };
        int best_len = 0, best_rounds = 0;
        for (int len = 4; len <= 4096; len++) {
            const int n = count(len);
            int r = 1;
            while (r * FMM_BLOCK < n) r *= 2;
            if (r > 8) continue;
            if (!best_len || r * len < best_rounds * best_len) {
                best_len = len;
                best_rounds = r;
            }
        }
        T.m2l_chunk.clear();
        T.m2l_cfirst.assign(K + 1, 0);
        for (int kb = 0; kb < K; kb++) {
            T.m2l_cfirst[kb] = (int)T.m2l_chunk.size() / 4;
            for (int t = T.m2l_off[kb]; t < T.m2l_off[kb + 1]; t += best_len) {
                T.m2l_chunk.push_back(kb);
                T.m2l_chunk.push_back(t);
                T.m2l_chunk.push_back(std::min(t + best_len, T.m2l_off[kb + 1]));
                T.m2l_chunk.push_back(0);
            }
        }
        T.m2l_nchunk = (int)T.m2l_chunk.size() / 4;
        T.m2l_cfirst[K] = T.m2l_nchunk;
        T.m2l_rounds = best_rounds;
        T.m2l_chunk_len = best_len;
    }
    // matrix-core M2L: scales and the E index of every tile element
    {
        auto fact = [](int n) {
            double f = 1.0;
            for (int i = 2; i <= n; i++) f *= i;
            return f;
        };
        T.m2l_bscale.assign(K, 0.0);
        T.m2l_lscale.assign(K, 0.0);
        for (int k = 0; k < K; k++) {
            const int *a = &M.mi[3 * k];
            const double f = fact(a[0]) * fact(a[1]) * fact(a[2]);
            T.m2l_bscale[k] = ((a[0] + a[1] + a[2]) % 2 ? -1.0 : 1.0) / f;
            T.m2l_lscale[k] = 1.0 / f;
        }
        T.m2l_erec.assign(4 * K, 0);
        for (int k = 0; k < K; k++) {
            int j[6];
            for (int i = 0; i < 3; i++) {
                j[i] = T.rec1[4 * k + i] >= 0 ? T.rec1[4 * k + i] : K;
                j[3 + i] = T.rec2[4 * k + i] >= 0 ? T.rec2[4 * k + i] : K;
            }
            // byte offsets (8 j) of the entries the recurrence reads, and
            // k! | |k| << 24 (k! <= 10! < 2^22 for the orders of the
            // matrix-core M2L)
            const int *m = &M.mi[3 * k];
            T.m2l_erec[4 * k] = 8 * j[0] | (8 * j[1] << 16);
            T.m2l_erec[4 * k + 1] = 8 * j[2] | (8 * j[3] << 16);
            T.m2l_erec[4 * k + 2] = 8 * j[4] | (8 * j[5] << 16);
            T.m2l_erec[4 * k + 3] =
                (int)(fact(m[0]) * fact(m[1]) * fact(m[2])) | ((m[0] + m[1] + m[2]) << 24);
        }
        T.m2l_tile_idx.clear();
        if (P >= 1 && P <= M2L_MFMA_MAX_P) {
            const int N8 = m2l_ntiles8(P);
            T.m2l_tile_idx.assign((size_t)16 * N8, (uint16_t)(8 * K));   // byte offsets
            int t = 0;
            for (int ta = 0; ta < m2l_ntile1(P); ta++)
                for (int tb = 0; tb < m2l_nrows(P, ta); tb++, t++)
                    for (int e = 0; e < 16; e++) {
                        const int row = e & 3, k = e >> 2, a = 4 * ta + k, b = 4 * tb + row;
                        if (a >= K || b >= K) continue;
                        const int *ma = &M.mi[3 * a], *mb = &M.mi[3 * b];
                        const int s = M.idx(ma[0] + mb[0], ma[1] + mb[1], ma[2] + mb[2]);
                        if (s >= 0) T.m2l_tile_idx[(size_t)e * N8 + t] = (uint16_t)(8 * s);
                    }
        }
    }
    // L2P: idx(b + e_j) and idx(b + e_j + e_l)
    T.d1.assign(4 * K, -1);
    T.d2a.assign(4 * K, -1);
    T.d2b.assign(4 * K, -1);
    const int pj[6] = {0, 0, 0, 1, 1, 2}, plx[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < K; k++) {
        const int *b = &M.mi[3 * k];
        for (int j = 0; j < 3; j++) {
            int e[3] = {b[0], b[1], b[2]};
            e[j]++;
            T.d1[4 * k + j] = M.idx(e[0], e[1], e[2]);
        }
        for (int m = 0; m < 6; m++) {
            int e[3] = {b[0], b[1], b[2]};
            e[pj[m]]++;
            e[plx[m]]++;
            const int v = M.idx(e[0], e[1], e[2]);
            if (m < 3) T.d2a[4 * k + m] = v; else T.d2b[4 * k + m - 3] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// SFS stretching over the near-field lists (Estr with the FMM)
// ---------------------------------------------------------------------------
// Estr(x_i) = sum_j zeta_sigma_j(x_i - x_j) (Gamma_j . grad)(U_i - U_j)
// (docs/src/theory/rvpm.md:251-262) decays like zeta: every pair the FMM
// treats as far field has a gap of at least sigcut * sigma_j, where zeta is
// below e^{-24.5} zeta(0) for the gaussian kernels, so the sum over the P2P
// lists is the whole sum to that precision.

// Estr store: the sums overwrite E (and with six sums, D[E] into DE).
template <int NV>
__device__ __forceinline__ void estr_finish(const double *d, int64_t i, int64_t cap,
                                            double *__restrict__ E, double *__restrict__ DE)
{
    E[i] = d[0];
    E[cap + i] = d[1];
    E[2 * cap + i] = d[2];
    if constexpr (NV == 6) {
        DE[i] = d[3];
        DE[cap + i] = d[4];
        DE[2 * cap + i] = d[5];
    }
}

// Sorted Estr source records (x, y, z, 1/sigma) (Gamma, 0) (J0..3) (J4..7)
// (J8, 0, 0, 0); with DYN the last is (J8, DJ0..2) followed by (DJ3..6)
// (DJ7, DJ8, 0, 0): 7 per particle.
template <bool DYN>
__global__ __launch_bounds__(FMM_BLOCK) void fmm_estr_gather_kernel(
    const double *__restrict__ X, const double *__restrict__ G, const double *__restrict__ sg,
    const double *__restrict__ Jm, const double *__restrict__ DJ,
    const int32_t *__restrict__ perm, int64_t n, int64_t cap, double4 *__restrict__ out)
{
    constexpr int NR = DYN ? 7 : 5;
    const int64_t s = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm[s];
    double4 *o = out + NR * s;
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

// The nine J (and with DYN DJ) entries of a record.
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

template <bool TRANSPOSED, typename T>
__device__ __forceinline__ void stretch_of(const T *dJ, T gx, T gy, T gz, T &S0, T &S1, T &S2)
{
    if constexpr (TRANSPOSED) {   // S_k = sum_l dJ[l,k] G_l
        S0 = dJ[0] * gx + dJ[1] * gy + dJ[2] * gz;
        S1 = dJ[3] * gx + dJ[4] * gy + dJ[5] * gz;
        S2 = dJ[6] * gx + dJ[7] * gy + dJ[8] * gz;
    } else {                      // S_k = sum_l dJ[k,l] G_l
        S0 = dJ[0] * gx + dJ[3] * gy + dJ[6] * gz;
        S1 = dJ[1] * gx + dJ[4] * gy + dJ[7] * gz;
        S2 = dJ[2] * gx + dJ[5] * gy + dJ[8] * gz;
    }
}

// Estr over the near-field lists, FP64; DYN adds D[E] (zeta_zD_f64).
template <int KK, bool TRANSPOSED, bool DYN>
__global__ __launch_bounds__(P2P_BLOCK) void fmm_estr_kernel(const P2PChunk *__restrict__ chunks,
                                                             const FmmCellDev *__restrict__ cells,
                                                             const int32_t *__restrict__ src,
                                                             const double4 *__restrict__ rec,
                                                             const int32_t *__restrict__ perm,
                                                             int64_t cap, double *__restrict__ E,
                                                             double *__restrict__ DE,
                                                             double *__restrict__ part,
                                                             int64_t pstride)
{
    constexpr int NR = DYN ? 7 : 5;
    constexpr int NV = DYN ? 6 : 3;
    __shared__ double4 tile[NR * P2P_BLOCK];
    const P2PChunk ch = chunks[blockIdx.x];
    const int c = ch.c;
    const FmmCellDev A = cells[c];
    for (int tb = A.begin; tb < A.end; tb += P2P_BLOCK) {
        const int it = tb + threadIdx.x;
        const bool active = it < A.end;
        double xi = 0, yi = 0, zi = 0, Ji[9], DJi[9];
        for (int k = 0; k < 9; ++k) Ji[k] = DJi[k] = 0.0;
        if (active) {
            const double4 p = rec[NR * it];
            xi = p.x; yi = p.y; zi = p.z;
            estr_rec_j<DYN>(rec + NR * it, Ji, DJi);
        }
        double e0 = 0, e1 = 0, e2 = 0, f0 = 0, f1 = 0, f2 = 0;
        for (int64_t e = ch.e0; e < ch.e1; e++) {
            const FmmCellDev B = cells[src[e]];
            for (int sb = B.begin; sb < B.end; sb += P2P_BLOCK) {
                __syncthreads();
                const int js = sb + threadIdx.x;
                if (js < B.end)
                    for (int k = 0; k < NR; ++k) tile[NR * threadIdx.x + k] = rec[NR * js + k];
                __syncthreads();
                const int cnt = min(P2P_BLOCK, B.end - sb);
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
                        zeta_zD_f64<KK>(r * a.w, z, zD);
                        const double zs = z * is3, zDs = zD * is3;
                        double T0, T1, T2;
                        for (int k = 0; k < 9; ++k) dJ[k] = DJi[k] - DJj[k];
                        stretch_of<TRANSPOSED>(dJ, gm.x, gm.y, gm.z, T0, T1, T2);
                        e0 += zs * S0; e1 += zs * S1; e2 += zs * S2;
                        f0 += zDs * S0 + zs * T0;
                        f1 += zDs * S1 + zs * T1;
                        f2 += zDs * S2 + zs * T2;
                    } else {
                        const double zs = zeta_f64<KK>(r * a.w) * is3;
                        e0 += zs * S0; e1 += zs * S1; e2 += zs * S2;
                    }
                }
            }
        }
        if (active) {
            const double d[6] = {e0, e1, e2, f0, f1, f2};
            if (ch.part < 0)
                estr_finish<NV>(d, perm[it], cap, E, DE);
            else
                for (int k = 0; k < NV; ++k) part[k * pstride + ch.part + (it - A.begin)] = d[k];
        }
    }
}

// Estr over the near-field lists, mixed: FP32 pair math on coordinates
// relative to the target leaf's centre and on J relative to the J of the
// leaf's first particle (J_i - J_j then keeps its relative precision), FP32
// partials per source tile folded into FP64.  Same streaming of the source
// leaves and replication of small target leaves as p2p_mixed_kernel.  DYN
// adds D[E] from DJ, taken relative to the first particle's DJ the same way.

// zeta(r / sigma) / sigma^3 and zD(r / sigma) / sigma^3 in FP32; p0 = 1/sigma,
// p1 = 1/sigma^2.
template <int KK, bool DYN>
__device__ __forceinline__ void zeta_s3_f32(float r2, float p0, float p1, float &zs, float &zDs)
{
    const float is3 = p0 * p1;
    if constexpr (KK == GAUSSIANERF) {
        zs = 0.06349363593424097f * fexp2(-0.5f * LOG2E * r2 * p1) * is3;   // (2 pi)^-3/2
        if constexpr (DYN) zDs = fmaf(r2, p1, -3.0f) * zs;                 // (rho^2 - 3) zeta
    } else if constexpr (KK == GAUSSIAN) {
        const float rho2 = r2 * p1, rho3 = rho2 * sqrtf(rho2);
        zs = 0.238732414637843f * fexp2(-LOG2E * rho3) * is3;               // 3/(4 pi)
        if constexpr (DYN) zDs = 3.0f * (rho3 - 1.0f) * zs;                 // 3 (rho^3 - 1) zeta
    } else {   // WINCKELMANS: 15/(8 pi) (rho^2 + 1)^-7/2
        const float si = frsq(fmaf(r2, p1, 1.0f)), si2 = si * si;
        zs = 0.5968310365946075f * si2 * si2 * si2 * si * is3;
        if constexpr (DYN) zDs = fmaf(-7.0f, si2, 4.0f) * zs;              // (7 rho^2/q - 3) zeta
    }
}

// LDS tile of the mixed Estr: per source position (x, y, z relative, 1/sigma),
// (Gamma, 1/sigma^2), J - Jr in 9 floats, and with DYN DJ - DJr in 9 floats.
struct EstrTile {
    float4 sp[P2P_BLOCK], sg[P2P_BLOCK], sj0[P2P_BLOCK], sj1[P2P_BLOCK];
    float sj8[P2P_BLOCK];
    float4 sd0[P2P_BLOCK], sd1[P2P_BLOCK];
    float sd8[P2P_BLOCK];
};

template <bool TRANSPOSED>
__device__ __forceinline__ void stretch_f32(float d0, float d1, float d2, float d3, float d4,
                                            float d5, float d6, float d7, float d8, const float4 g,
                                            float &S0, float &S1, float &S2)
{
    if constexpr (TRANSPOSED) {
        S0 = fmaf(d0, g.x, fmaf(d1, g.y, d2 * g.z));
        S1 = fmaf(d3, g.x, fmaf(d4, g.y, d5 * g.z));
        S2 = fmaf(d6, g.x, fmaf(d7, g.y, d8 * g.z));
    } else {
        S0 = fmaf(d0, g.x, fmaf(d3, g.y, d6 * g.z));
        S1 = fmaf(d1, g.x, fmaf(d4, g.y, d7 * g.z));
        S2 = fmaf(d2, g.x, fmaf(d5, g.y, d8 * g.z));
    }
}

template <int KK, bool TRANSPOSED, bool DYN>
__device__ __forceinline__ void estr_mixed_pair(const EstrTile &T, int j, float xi, float yi,
                                                float zi, const float (&Ji)[9],
                                                const float (&DJi)[9], float (&e)[6])
{
    const float4 a = T.sp[j], g = T.sg[j], q0 = T.sj0[j], q1 = T.sj1[j];
    const float dx = xi - a.x, dy = yi - a.y, dz = zi - a.z;
    float zs, zDs = 0.f;
    zeta_s3_f32<KK, DYN>(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), a.w, g.w, zs, zDs);
    float S0, S1, S2;
    stretch_f32<TRANSPOSED>(Ji[0] - q0.x, Ji[1] - q0.y, Ji[2] - q0.z, Ji[3] - q0.w, Ji[4] - q1.x,
                            Ji[5] - q1.y, Ji[6] - q1.z, Ji[7] - q1.w, Ji[8] - T.sj8[j], g, S0, S1,
                            S2);
    e[0] = fmaf(zs, S0, e[0]);
    e[1] = fmaf(zs, S1, e[1]);
    e[2] = fmaf(zs, S2, e[2]);
    if constexpr (DYN) {
        const float4 r0 = T.sd0[j], r1 = T.sd1[j];
        float T0, T1, T2;
        stretch_f32<TRANSPOSED>(DJi[0] - r0.x, DJi[1] - r0.y, DJi[2] - r0.z, DJi[3] - r0.w,
                                DJi[4] - r1.x, DJi[5] - r1.y, DJi[6] - r1.z, DJi[7] - r1.w,
                                DJi[8] - T.sd8[j], g, T0, T1, T2);
        e[3] = fmaf(zDs, S0, fmaf(zs, T0, e[3]));
        e[4] = fmaf(zDs, S1, fmaf(zs, T1, e[4]));
        e[5] = fmaf(zDs, S2, fmaf(zs, T2, e[5]));
    }
}

template <int KK, bool TRANSPOSED, bool DYN, int R>
__device__ __forceinline__ void estr_mixed_tile(const EstrTile &T, int cnt, int s, float xi,
                                                float yi, float zi, const float (&Ji)[9],
                                                const float (&DJi)[9], float (&e)[6])
{
    if (cnt == P2P_BLOCK) {
#pragma unroll 4
        for (int i = 0; i < P2P_BLOCK / R; ++i)
            estr_mixed_pair<KK, TRANSPOSED, DYN>(T, i * R + s, xi, yi, zi, Ji, DJi, e);
    } else {
        for (int j = s; j < cnt; j += R)
            estr_mixed_pair<KK, TRANSPOSED, DYN>(T, j, xi, yi, zi, Ji, DJi, e);
    }
}

template <int KK, bool TRANSPOSED, bool DYN>
__global__ __launch_bounds__(P2P_BLOCK) void fmm_estr_mixed_kernel(
    const P2PChunk *__restrict__ chunks, const FmmCellDev *__restrict__ cells,
    const int32_t *__restrict__ src, const double4 *__restrict__ rec,
    const int32_t *__restrict__ perm, int64_t cap, double *__restrict__ E,
    double *__restrict__ DE, double *__restrict__ part, int64_t pstride)
{
    constexpr int NR = DYN ? 7 : 5;
    constexpr int NV = DYN ? 6 : 3;
    __shared__ EstrTile T;
    __shared__ int32_t lbeg[P2P_SLOTS], lpre[P2P_SLOTS + 1];
    __shared__ double red[NV][P2P_BLOCK];
    const int lane = threadIdx.x;
    const P2PChunk ch = chunks[blockIdx.x];
    const int c = ch.c;
    const FmmCellDev A = cells[c];
    const int64_t e0 = ch.e0, e1 = ch.e1;
    for (int tb = A.begin; tb < A.end; tb += P2P_BLOCK) {
        const int nt = min(P2P_BLOCK, A.end - tb);
        const int R = nt <= 8 ? 8 : nt <= 16 ? 4 : nt <= 32 ? 2 : 1;
        const int sl = lane / nt;                  // slice of this lane
        const int t = lane - sl * nt;              // target of this lane
        const bool active = sl < R;
        // reference J (and DJ): the block's first target
        double Jr[9], DJr[9];
        for (int k = 0; k < 9; ++k) DJr[k] = 0.0;
        estr_rec_j<DYN>(rec + NR * tb, Jr, DJr);
        float xi, yi, zi, Ji[9], DJi[9];
        {
            double Jt[9], DJt[9];
            for (int k = 0; k < 9; ++k) DJt[k] = 0.0;
            const double4 a = rec[NR * (tb + t)];
            estr_rec_j<DYN>(rec + NR * (tb + t), Jt, DJt);
            xi = (float)(a.x - A.cx); yi = (float)(a.y - A.cy); zi = (float)(a.z - A.cz);
            for (int k = 0; k < 9; ++k) {
                Ji[k] = (float)(Jt[k] - Jr[k]);
                DJi[k] = (float)(DJt[k] - DJr[k]);
            }
        }
        const int s = active ? sl : 0;
        double d[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t eb = e0; eb < e1; eb += P2P_SLOTS) {
            const int nl = (int)min<int64_t>(P2P_SLOTS, e1 - eb);
            __syncthreads();
            const int total = stream_setup(cells, src, eb, nl, lane, lbeg, lpre);
            // the records of the next tile load while this one computes
            double4 rr[NR];
            for (int k = 0; k < NR; ++k) rr[k] = make_double4(0.0, 0.0, 0.0, 0.0);
            if (lane < total) {
                const int js = stream_particle(lbeg, lpre, nl, lane);
                for (int k = 0; k < NR; ++k) rr[k] = rec[NR * js + k];
            }
            for (int base = 0; base < total; base += P2P_BLOCK) {
                __syncthreads();
                if (base + lane < total) {
                    const double4 ra = rr[0], rg = rr[1], ru = rr[2], rv = rr[3], rw = rr[4];
                    T.sp[lane] = make_float4((float)(ra.x - A.cx), (float)(ra.y - A.cy),
                                             (float)(ra.z - A.cz), (float)ra.w);
                    T.sg[lane] = make_float4((float)rg.x, (float)rg.y, (float)rg.z,
                                             (float)(ra.w * ra.w));
                    T.sj0[lane] = make_float4((float)(ru.x - Jr[0]), (float)(ru.y - Jr[1]),
                                              (float)(ru.z - Jr[2]), (float)(ru.w - Jr[3]));
                    T.sj1[lane] = make_float4((float)(rv.x - Jr[4]), (float)(rv.y - Jr[5]),
                                              (float)(rv.z - Jr[6]), (float)(rv.w - Jr[7]));
                    T.sj8[lane] = (float)(rw.x - Jr[8]);
                    if constexpr (DYN) {
                        const double4 rx = rr[5], ry = rr[6];
                        T.sd0[lane] = make_float4((float)(rw.y - DJr[0]), (float)(rw.z - DJr[1]),
                                                  (float)(rw.w - DJr[2]), (float)(rx.x - DJr[3]));
                        T.sd1[lane] = make_float4((float)(rx.y - DJr[4]), (float)(rx.z - DJr[5]),
                                                  (float)(rx.w - DJr[6]), (float)(ry.x - DJr[7]));
                        T.sd8[lane] = (float)(ry.y - DJr[8]);
                    }
                }
                __syncthreads();
                const int qn = base + P2P_BLOCK + lane;
                if (qn < total) {
                    const int js = stream_particle(lbeg, lpre, nl, qn);
                    for (int k = 0; k < NR; ++k) rr[k] = rec[NR * js + k];
                }
                const int cnt = min(P2P_BLOCK, total - base);
                float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                switch (R) {
                case 8: estr_mixed_tile<KK, TRANSPOSED, DYN, 8>(T, cnt, s, xi, yi, zi, Ji, DJi, f); break;
                case 4: estr_mixed_tile<KK, TRANSPOSED, DYN, 4>(T, cnt, s, xi, yi, zi, Ji, DJi, f); break;
                case 2: estr_mixed_tile<KK, TRANSPOSED, DYN, 2>(T, cnt, s, xi, yi, zi, Ji, DJi, f); break;
                default: estr_mixed_tile<KK, TRANSPOSED, DYN, 1>(T, cnt, s, xi, yi, zi, Ji, DJi, f); break;
                }
                for (int k = 0; k < NV; ++k) d[k] += f[k];
            }
        }
        if (R > 1) {   // add the slices of each target in slice order
            __syncthreads();
            for (int k = 0; k < NV; ++k) red[k][lane] = d[k];
            __syncthreads();
            if (lane < nt) {
                for (int q = 1; q < R; ++q)
                    for (int k = 0; k < NV; ++k) d[k] += red[k][lane + q * nt];
            }
        }
        if (lane < nt) {
            if (ch.part < 0)
                estr_finish<NV>(d, perm[tb + lane], cap, E, DE);
            else
                for (int k = 0; k < NV; ++k)
                    part[k * pstride + ch.part + (tb + lane - A.begin)] = d[k];
        }
    }
}

// Partial sums of the leaves cut into several work items, added in item order
// and stored.  MODE 0: FP64 P2P (12 sums), 1: mixed P2P (14), 2: Estr (3),
// 3: Estr and D[E] (6).
template <int MODE>
__global__ __launch_bounds__(P2P_BLOCK) void p2p_part_reduce_kernel(
    const int32_t *__restrict__ leaves, const int64_t *__restrict__ coff,
    const int64_t *__restrict__ pbase, const FmmCellDev *__restrict__ cells,
    const double *__restrict__ part, int64_t pstride, const int32_t *__restrict__ perm,
    int64_t cap, double *__restrict__ U, double *__restrict__ J, double *__restrict__ E)
{
    constexpr int NV = MODE == 0 ? 12 : MODE == 1 ? 14 : MODE == 2 ? 3 : 6;
    const int64_t li = blockIdx.x;
    const int64_t n = coff[li + 1] - coff[li];
    if (n <= 1) return;
    const FmmCellDev A = cells[leaves[li]];
    const int nt = A.end - A.begin;
    const int64_t pb = pbase[li];
    for (int t = threadIdx.x; t < nt; t += P2P_BLOCK) {
        double d[NV];
        for (int k = 0; k < NV; ++k) d[k] = 0.0;
        for (int64_t j = 0; j < n; ++j)
            for (int k = 0; k < NV; ++k) d[k] += part[k * pstride + pb + j * nt + t];
        const int64_t i = perm[A.begin + t];
        if constexpr (MODE == 0)
            p2p_f64_finish(d, i, cap, U, J);
        else if constexpr (MODE == 1)
            p2p_mixed_finish(d, i, cap, U, J);
        else if constexpr (MODE == 2)
            estr_finish<3>(d, i, cap, E, nullptr);
        else
            estr_finish<6>(d, i, cap, E, U);   // U carries DE here
    }
}

template <int MODE>
static hipError_t launch_part_reduce(hipStream_t st, const int32_t *leaves, int64_t nleaf,
                                     const int64_t *coff, const int64_t *pbase,
                                     const FmmCellDev *cells, const double *part, int64_t nslot,
                                     const int32_t *perm, int64_t cap, double *U, double *J,
                                     double *E)
{
    if (nslot == 0 || nleaf == 0) return hipSuccess;
    p2p_part_reduce_kernel<MODE><<<(unsigned)nleaf, P2P_BLOCK, 0, st>>>(
        leaves, coff, pbase, cells, part, nslot, perm, cap, U, J, E);
    return hipGetLastError();
}

hipError_t launch_fmm_estr(hipStream_t st, const FmmState &S, int kernel, int transposed,
                           int precision, const DeviceField &f, Scratch &s, const double *DJ,
                           double *DE)
{
    if (f.np == 0) return hipSuccess;
    const bool dyn = DJ != nullptr;
    if (dyn)
        fmm_estr_gather_kernel<true><<<gridn(f.np, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
            f.X, f.G, f.sigma, f.J, DJ, s.perm, f.np, f.cap, s.srcE);
    else
        fmm_estr_gather_kernel<false><<<gridn(f.np, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
            f.X, f.G, f.sigma, f.J, DJ, s.perm, f.np, f.cap, s.srcE);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nleaf = (int64_t)S.plan.nleaves;
    if (nleaf == 0 || S.nchunk == 0) return hipSuccess;
    const unsigned g = (unsigned)S.nchunk;
#define VPMX_FESTR3(KK, TR, DY)                                                                  \
    if (precision)                                                                               \
        fmm_estr_mixed_kernel<KK, TR, DY><<<g, P2P_BLOCK, 0, st>>>(                             \
            S.chunks, S.cells, S.p2p_src, s.srcE, s.perm, f.cap, f.E, DE, S.part, S.nslot);      \
    else                                                                                         \
        fmm_estr_kernel<KK, TR, DY><<<g, P2P_BLOCK, 0, st>>>(S.chunks, S.cells, S.p2p_src,      \
                                                            s.srcE, s.perm, f.cap, f.E, DE,      \
                                                            S.part, S.nslot);
#define VPMX_FESTR(KK)                                                                           \
    if (dyn && transposed) {                                                                     \
        VPMX_FESTR3(KK, true, true)                                                              \
    } else if (dyn) {                                                                            \
        VPMX_FESTR3(KK, false, true)                                                             \
    } else if (transposed) {                                                                     \
        VPMX_FESTR3(KK, true, false)                                                             \
    } else {                                                                                     \
        VPMX_FESTR3(KK, false, false)                                                            \
    }
    switch (kernel) {
    case GAUSSIAN: VPMX_FESTR(GAUSSIAN); break;
    case GAUSSIANERF: VPMX_FESTR(GAUSSIANERF); break;
    case WINCKELMANS: VPMX_FESTR(WINCKELMANS); break;
    default: return hipErrorInvalidValue;
    }
#undef VPMX_FESTR
#undef VPMX_FESTR3
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (dyn)
        return launch_part_reduce<3>(st, S.leaves, nleaf, S.chunk_off, S.part_base, S.cells,
                                     S.part, S.nslot, s.perm, f.cap, DE, nullptr, f.E);
    return launch_part_reduce<2>(st, S.leaves, nleaf, S.chunk_off, S.part_base, S.cells, S.part,
                                 S.nslot, s.perm, f.cap, nullptr, nullptr, f.E);
}

// Core-size derivatives over the near-field lists: the P2P kernels in their
// DS mode, adding into DU, DJ cleared first.
__global__ __launch_bounds__(FMM_BLOCK) void fmm_zero_kernel(double *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

hipError_t launch_fmm_dsig(hipStream_t st, const FmmState &S, int kernel, int precision,
                           const DeviceField &f, const Scratch &s, double *DU, double *DJ)
{
    if (f.np == 0) return hipSuccess;
    fmm_zero_kernel<<<gridn(3 * f.cap, FMM_BLOCK), FMM_BLOCK, 0, st>>>(DU, 3 * f.cap);
    fmm_zero_kernel<<<gridn(9 * f.cap, FMM_BLOCK), FMM_BLOCK, 0, st>>>(DJ, 9 * f.cap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nleaf = (int64_t)S.plan.nleaves;
    if (nleaf == 0 || S.nchunk == 0) return hipSuccess;
    const unsigned g = (unsigned)S.nchunk;
#define VPMX_DS(KK)                                                                             \
    if (precision == 0)                                                                         \
        p2p_f64_kernel<KK, true><<<g, P2P_BLOCK, 0, st>>>(S.chunks, S.cells, S.p2p_src, s.src64, \
                                                         s.perm, f.cap, DU, DJ, S.part, S.nslot); \
    else                                                                                        \
        p2p_mixed_kernel<KK, true><<<g, P2P_BLOCK, 0, st>>>(S.chunks, S.cells, S.p2p_src,       \
                                                           s.src64, s.perm, f.cap, DU, DJ,       \
                                                           S.part, S.nslot);
    switch (kernel) {
    case GAUSSIAN: VPMX_DS(GAUSSIAN); break;
    case GAUSSIANERF: VPMX_DS(GAUSSIANERF); break;
    case WINCKELMANS: VPMX_DS(WINCKELMANS); break;
    default: return hipErrorInvalidValue;
    }
#undef VPMX_DS
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return precision == 0
               ? launch_part_reduce<0>(st, S.leaves, nleaf, S.chunk_off, S.part_base, S.cells,
                                       S.part, S.nslot, s.perm, f.cap, DU, DJ, nullptr)
               : launch_part_reduce<1>(st, S.leaves, nleaf, S.chunk_off, S.part_base, S.cells,
                                       S.part, S.nslot, s.perm, f.cap, DU, DJ, nullptr);
}

// ---------------------------------------------------------------------------
// Viscous sums over the near-field work items (FP64)
// ---------------------------------------------------------------------------
// The FMM leaves zeta (and PSE's eta = -zeta'/rho, both decaying within a few
// cores for the gaussian kernels) to its near field, as FLOWVPM's zeta_fmm
// does (examples/rotorhover/rotorhover.jl:170): these run over the work items
// of the last FMM evaluation's P2P lists, on records in its sorted order.
//   PSE = false: w_i = sum_j Gamma_j zeta(r / s_j) / s_j^3 (coincident pairs
//                included); records (x, y, z, 1/s) (Gx, Gy, Gz, 1/s^3), masked
//                sources with no strength.
//   PSE = true:  V_i = sum_j (2 nu / s^2) (vol_i Gamma_j - vol_j Gamma_i)
//                eta(r / s) / s^3, s^2 = (sigma_i^2 + sigma_j^2) / 2, j != i;
//                records (x, y, z, sigma^2) (Gx, Gy, Gz, vol).
template <bool PSE>
__global__ __launch_bounds__(P2P_BLOCK) void pack_visc_kernel(
    const double *__restrict__ X, const double *__restrict__ G, int64_t gld,
    const double *__restrict__ sg, double s0, const double *__restrict__ vol,
    const uint8_t *__restrict__ mask, const int32_t *__restrict__ perm, int64_t n, int64_t cap,
    double4 *__restrict__ rec)
{
    const int64_t s = (int64_t)blockIdx.x * P2P_BLOCK + threadIdx.x;
    if (s >= n) return;
    const int64_t i = perm[s];
    const double x = X[i], y = X[cap + i], z = X[2 * cap + i];
    if constexpr (PSE) {
        const double sig = sg[i];
        rec[2 * s] = make_double4(x, y, z, sig * sig);
        rec[2 * s + 1] = make_double4(G[i], G[gld + i], G[2 * gld + i], vol[i]);
    } else {
        const double is = 1.0 / (s0 > 0.0 ? s0 : sg[i]);
        const bool off = mask && mask[i];
        rec[2 * s] = make_double4(x, y, z, is);
        rec[2 * s + 1] = make_double4(off ? 0.0 : G[i], off ? 0.0 : G[gld + i],
                                      off ? 0.0 : G[2 * gld + i], is * is * is);
    }
}

template <int KK, bool PSE>
__global__ __launch_bounds__(P2P_BLOCK) void near_visc_kernel(
    const P2PChunk *__restrict__ chunks, const FmmCellDev *__restrict__ cells,
    const int32_t *__restrict__ src, const double4 *__restrict__ rec,
    const int32_t *__restrict__ perm, double nu, int64_t ld, double *__restrict__ out,
    double *__restrict__ part, int64_t pstride)
{
    __shared__ double4 tile[2 * P2P_BLOCK];
    const P2PChunk ch = chunks[blockIdx.x];
    const FmmCellDev A = cells[ch.c];
    for (int tb = A.begin; tb < A.end; tb += P2P_BLOCK) {
        const int it = tb + threadIdx.x;
        const bool active = it < A.end;
        double4 a = make_double4(0.0, 0.0, 0.0, 1.0), b = make_double4(0.0, 0.0, 0.0, 0.0);
        if (active) {
            a = rec[2 * it];
            b = rec[2 * it + 1];
        }
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        for (int64_t e = ch.e0; e < ch.e1; e++) {
            const FmmCellDev B = cells[src[e]];
            for (int sb = B.begin; sb < B.end; sb += P2P_BLOCK) {
                __syncthreads();
                const int js = sb + threadIdx.x;
                if (js < B.end) {
                    tile[2 * threadIdx.x] = rec[2 * js];
                    tile[2 * threadIdx.x + 1] = rec[2 * js + 1];
                }
                __syncthreads();
                const int cnt = min(P2P_BLOCK, B.end - sb);
                for (int jj = 0; jj < cnt; ++jj) {
                    const double4 p = tile[2 * jj], q = tile[2 * jj + 1];
                    const double d0 = a.x - p.x, d1 = a.y - p.y, d2 = a.z - p.z;
                    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    if constexpr (PSE) {
                        if (sb + jj == it) continue;
                        const double s2 = 0.5 * (a.w + p.w);
                        const double sq = sqrt(s2);
                        const double w = 2.0 * nu / s2 * eta_f64<KK>(r / sq) / (s2 * sq);
                        v0 += w * (b.w * q.x - q.w * b.x);
                        v1 += w * (b.w * q.y - q.w * b.y);
                        v2 += w * (b.w * q.z - q.w * b.z);
                    } else {
                        const double z = zeta_f64<KK>(r * p.w) * q.w;
                        v0 += z * q.x;
                        v1 += z * q.y;
                        v2 += z * q.z;
                    }
                }
            }
        }
        if (active) {
            if (ch.part < 0) {
                const int64_t i = perm[it];
                out[i] = v0;
                out[ld + i] = v1;
                out[2 * ld + i] = v2;
            } else {
                const int64_t o = ch.part + (it - A.begin);
                part[o] = v0;
                part[pstride + o] = v1;
                part[2 * pstride + o] = v2;
            }
        }
    }
}

template <bool PSE>
static hipError_t near_visc_go(hipStream_t st, const FmmState &S, int kernel,
                               const DeviceField &f, const Scratch &s, const double *G,
                               int64_t gld, double s0, const uint8_t *mask, double nu,
                               double4 *rec, double *out, int64_t ld)
{
    const int64_t n = f.np;
    if (n == 0) return hipSuccess;
    pack_visc_kernel<PSE><<<gridn(n, P2P_BLOCK), P2P_BLOCK, 0, st>>>(
        f.X, G, gld, f.sigma, s0, f.vol, mask, s.perm, n, f.cap, rec);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nleaf = (int64_t)S.plan.nleaves;
    if (nleaf == 0 || S.nchunk == 0) return hipSuccess;
    const unsigned g = (unsigned)S.nchunk;
    switch (kernel) {
    case GAUSSIAN:
        near_visc_kernel<GAUSSIAN, PSE><<<g, P2P_BLOCK, 0, st>>>(
            S.chunks, S.cells, S.p2p_src, rec, s.perm, nu, ld, out, S.part, S.nslot);
        break;
    case GAUSSIANERF:
        near_visc_kernel<GAUSSIANERF, PSE><<<g, P2P_BLOCK, 0, st>>>(
            S.chunks, S.cells, S.p2p_src, rec, s.perm, nu, ld, out, S.part, S.nslot);
        break;
    case WINCKELMANS:
        near_visc_kernel<WINCKELMANS, PSE><<<g, P2P_BLOCK, 0, st>>>(
            S.chunks, S.cells, S.p2p_src, rec, s.perm, nu, ld, out, S.part, S.nslot);
        break;
    default:
        return hipErrorInvalidValue;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // leaves cut into several items: their partials added in item order
    return launch_part_reduce<2>(st, S.leaves, nleaf, S.chunk_off, S.part_base, S.cells, S.part,
                                 S.nslot, s.perm, ld, nullptr, nullptr, out);
}

hipError_t launch_fmm_zeta(hipStream_t st, const FmmState &S, int kernel, const DeviceField &f,
                           const Scratch &s, const double *G, int64_t gld, double s0,
                           const uint8_t *mask, double4 *rec, double *w, int64_t wld)
{
    return near_visc_go<false>(st, S, kernel, f, s, G, gld, s0, mask, 0.0, rec, w, wld);
}

hipError_t launch_fmm_pse(hipStream_t st, const FmmState &S, int kernel, const DeviceField &f,
                          const Scratch &s, double nu, double4 *rec)
{
    return near_visc_go<true>(st, S, kernel, f, s, f.G, f.cap, 0.0, nullptr, nu, rec, f.V, f.cap);
}

// ---------------------------------------------------------------------------
// Device dual tree traversal
// ---------------------------------------------------------------------------
// Level-synchronous: every frontier pair (target a, source b) is classified by
// the same rules as fmm_traverse (M2L, P2P, or a split into child pairs);
// exclusive scans of the per-pair output counts give each pair fixed output
// slots, so the lists are deterministic without atomics.  The M2L and P2P
// lists are then radix-sorted (stable) on the target bits of their
// (target, source) keys only: grouped by target for the CSR, each target's
// sources in emission order (the host traversal orders them by source; the
// lists hold the same pairs).

// Multipole acceptance with IEEE-rounded operations in the host's order (no
// FMA contraction), so device and host decisions agree bit for bit.
__device__ __forceinline__ bool mac_dev(const FmmCellDev &A, const FmmCellDev &B, double theta,
                                        double sigcut)
{
#pragma clang fp contract(off)
    const double dx = A.cx - B.cx, dy = A.cy - B.cy, dz = A.cz - B.cz;
    const double d = __dsqrt_rn(dx * dx + dy * dy + dz * dz);
    return (A.R + B.R) < theta * d && (d - A.R - B.R) >= sigcut * B.smax;
}

// kind: 0 M2L, 1 P2P, 2 split both (a == b), 3 split target, 4 split source
__device__ __forceinline__ int dtt_action(const int2 pr, const FmmCellDev *__restrict__ cells,
                                          double theta, double sigcut, int &nsplit)
{
    const FmmCellDev A = cells[pr.x], B = cells[pr.y];
    nsplit = 0;
    if (pr.x == pr.y) {
        if (A.nchild == 0) return 1;
        nsplit = A.nchild * A.nchild;
        return 2;
    }
    if (mac_dev(A, B, theta, sigcut)) return 0;
    if (A.nchild == 0 && B.nchild == 0) return 1;
    if (B.nchild == 0 || (A.nchild != 0 && A.R >= B.R)) {
        nsplit = A.nchild;
        return 3;
    }
    nsplit = B.nchild;
    return 4;
}

__global__ __launch_bounds__(FMM_BLOCK) void dtt_classify_kernel(const int2 *__restrict__ fr,
                                                                 int64_t nf,
                                                                 const FmmCellDev *__restrict__ cells,
                                                                 double theta, double sigcut,
                                                                 int *__restrict__ cnext,
                                                                 int *__restrict__ cm2l,
                                                                 int *__restrict__ cp2p)
{
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i >= nf) return;
    int ns;
    const int kind = dtt_action(fr[i], cells, theta, sigcut, ns);
    cnext[i] = ns;
    cm2l[i] = kind == 0;
    cp2p[i] = kind == 1;
}

__global__ __launch_bounds__(FMM_BLOCK) void dtt_emit_kernel(
    const int2 *__restrict__ fr, int64_t nf, const FmmCellDev *__restrict__ cells, double theta,
    double sigcut, const int64_t *__restrict__ onext, const int64_t *__restrict__ om2l,
    const int64_t *__restrict__ op2p, int2 *__restrict__ next, uint64_t *__restrict__ m2l,
    uint64_t *__restrict__ p2p, int64_t m2l_base, int64_t p2p_base, int nb)
{
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i >= nf) return;
    const int2 pr = fr[i];
    int ns;
    const int kind = dtt_action(pr, cells, theta, sigcut, ns);
    const uint64_t key = ((uint64_t)pr.x << nb) | (uint64_t)pr.y;
    if (kind == 0) {
        m2l[m2l_base + om2l[i]] = key;
    } else if (kind == 1) {
        p2p[p2p_base + op2p[i]] = key;
    } else {
        const FmmCellDev A = cells[pr.x], B = cells[pr.y];
        int64_t o = onext[i];
        if (kind == 2) {
            for (int a = 0; a < A.nchild; a++)
                for (int b = 0; b < A.nchild; b++) next[o++] = make_int2(A.child0 + a, A.child0 + b);
        } else if (kind == 3) {
            for (int a = 0; a < A.nchild; a++) next[o++] = make_int2(A.child0 + a, pr.y);
        } else {
            for (int b = 0; b < B.nchild; b++) next[o++] = make_int2(pr.x, B.child0 + b);
        }
    }
}

// totals[k] = off_k[n-1] + cnt_k[n-1] for the three output kinds
__global__ void dtt_totals_kernel(int64_t n, const int *cnext, const int *cm2l, const int *cp2p,
                                  const int64_t *onext, const int64_t *om2l, const int64_t *op2p,
                                  int64_t *totals)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    totals[0] = onext[n - 1] + cnext[n - 1];
    totals[1] = om2l[n - 1] + cm2l[n - 1];
    totals[2] = op2p[n - 1] + cp2p[n - 1];
}

// CSR offsets from keys sorted by target: off[c] = first key with target >= c.
__global__ __launch_bounds__(FMM_BLOCK) void csr_kernel(const uint64_t *__restrict__ keys,
                                                        int64_t nkeys, int64_t ncell, int nb,
                                                        int64_t *__restrict__ off,
                                                        int32_t *__restrict__ src)
{
    const int64_t t = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (t <= ncell) {
        const uint64_t v = (uint64_t)t << nb;
        int64_t lo = 0, hi = nkeys;
        while (lo < hi) {
            const int64_t m = lo + (hi - lo) / 2;
            if (keys[m] < v) lo = m + 1; else hi = m;
        }
        off[t] = lo;
    }
    const uint64_t mask = ((uint64_t)1 << nb) - 1;
    for (int64_t i = t; i < nkeys; i += (int64_t)gridDim.x * FMM_BLOCK) src[i] = (int32_t)(keys[i] & mask);
}

// Particle pairs of the P2P lists, one thread per sorted (target, source) key,
// a wave-then-block sum and one integer atomic per block (order-independent).
__global__ __launch_bounds__(FMM_BLOCK) void p2p_count_kernel(const uint64_t *__restrict__ keys,
                                                              int64_t n, int nb,
                                                              const FmmCellDev *__restrict__ cells,
                                                              unsigned long long *__restrict__ total)
{
    __shared__ unsigned long long part[FMM_BLOCK / 64];
    const uint64_t mask = ((uint64_t)1 << nb) - 1;
    unsigned long long s = 0;
    for (int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * FMM_BLOCK) {
        const uint64_t k = keys[i];
        const FmmCellDev A = cells[k >> nb], B = cells[k & mask];
        s += (unsigned long long)(A.end - A.begin) * (unsigned long long)(B.end - B.begin);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < FMM_BLOCK / 64; ++w) t += part[w];
        if (t) atomicAdd(total, t);
    }
}

template <class T>
static hipError_t grow(T **p, int64_t &cap, int64_t need, hipStream_t st, int64_t keep = 0)
{
    if (need <= cap && *p) return hipSuccess;
    const int64_t nc = std::max<int64_t>({need, cap + cap / 2, 1024});
    T *q = nullptr;
    hipError_t e = hipMalloc((void **)&q, nc * sizeof(T));
    if (e != hipSuccess) return e;
    if (*p && keep > 0) {
        e = hipMemcpyAsync(q, *p, keep * sizeof(T), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    if (*p) (void)hipFree(*p);
    *p = q;
    cap = nc;
    return hipSuccess;
}

#define VPMX_CHK(x)                              \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return e_;         \
    } while (0)

hipError_t fmm_traverse_device(hipStream_t st, FmmState &S, int64_t ncell, double theta,
                               double sigcut, DttResult &res)
{
    DttBufs &B = S.dtt;
    res = DttResult();
    int nb = 1;
    while (((int64_t)1 << nb) <= ncell) nb++;
    if (!B.totals) {
        VPMX_CHK(hipMalloc((void **)&B.totals, 4 * sizeof(int64_t)));
        VPMX_CHK(hipHostMalloc((void **)&B.h_totals, 4 * sizeof(int64_t), hipHostMallocDefault));
    }
    VPMX_CHK(grow(&B.front, B.cap_front, 1, st));
    VPMX_CHK(grow(&B.next, B.cap_next, 1, st));
    const int2 root = make_int2(0, 0);
    VPMX_CHK(hipMemcpyAsync(B.front, &root, sizeof(int2), hipMemcpyHostToDevice, st));
    int64_t nf = 1, nm2l = 0, np2p = 0;
    while (nf > 0) {
        res.steps++;
        if (nf > B.cap_count) {
            int64_t c1 = B.cap_count, c2 = B.cap_count, c3 = B.cap_count, c4 = B.cap_count,
                    c5 = B.cap_count, c6 = B.cap_count;
            VPMX_CHK(grow(&B.cnt_next, c1, nf, st));
            VPMX_CHK(grow(&B.cnt_m2l, c2, nf, st));
            VPMX_CHK(grow(&B.cnt_p2p, c3, nf, st));
            VPMX_CHK(grow(&B.off_next, c4, nf, st));
            VPMX_CHK(grow(&B.off_m2l, c5, nf, st));
            VPMX_CHK(grow(&B.off_p2p, c6, nf, st));
            B.cap_count = c1;
            size_t need = 0;
            VPMX_CHK(rocprim::exclusive_scan(nullptr, need, B.cnt_next, B.off_next, (int64_t)0,
                                             (size_t)B.cap_count, rocprim::plus<int64_t>(), st));
            if (need > B.tmp_bytes) {
                if (B.tmp) (void)hipFree(B.tmp);
                B.tmp = nullptr;
                VPMX_CHK(hipMalloc((void **)&B.tmp, need));
                B.tmp_bytes = need;
            }
        }
        dtt_classify_kernel<<<gridn(nf, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
            B.front, nf, S.cells, theta, sigcut, B.cnt_next, B.cnt_m2l, B.cnt_p2p);
        VPMX_CHK(hipGetLastError());
        size_t tb = B.tmp_bytes;
        VPMX_CHK(rocprim::exclusive_scan(B.tmp, tb, B.cnt_next, B.off_next, (int64_t)0, (size_t)nf,
                                         rocprim::plus<int64_t>(), st));
        tb = B.tmp_bytes;
        VPMX_CHK(rocprim::exclusive_scan(B.tmp, tb, B.cnt_m2l, B.off_m2l, (int64_t)0, (size_t)nf,
                                         rocprim::plus<int64_t>(), st));
        tb = B.tmp_bytes;
        VPMX_CHK(rocprim::exclusive_scan(B.tmp, tb, B.cnt_p2p, B.off_p2p, (int64_t)0, (size_t)nf,
                                         rocprim::plus<int64_t>(), st));
        dtt_totals_kernel<<<1, 64, 0, st>>>(nf, B.cnt_next, B.cnt_m2l, B.cnt_p2p, B.off_next,
                                            B.off_m2l, B.off_p2p, B.totals);
        VPMX_CHK(hipGetLastError());
        VPMX_CHK(hipMemcpyAsync(B.h_totals, B.totals, 3 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        VPMX_CHK(hipStreamSynchronize(st));
        const int64_t nn = B.h_totals[0], am = B.h_totals[1], ap = B.h_totals[2];
        VPMX_CHK(grow(&B.next, B.cap_next, nn, st));
        VPMX_CHK(grow(&B.m2l_keys, B.cap_m2l, nm2l + am, st, nm2l));
        VPMX_CHK(grow(&B.p2p_keys, B.cap_p2p, np2p + ap, st, np2p));
        dtt_emit_kernel<<<gridn(nf, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
            B.front, nf, S.cells, theta, sigcut, B.off_next, B.off_m2l, B.off_p2p, B.next,
            B.m2l_keys, B.p2p_keys, nm2l, np2p, nb);
        VPMX_CHK(hipGetLastError());
        nm2l += am;
        np2p += ap;
        std::swap(B.front, B.next);
        std::swap(B.cap_front, B.cap_next);
        nf = nn;
    }
    // sort both lists by (target, source) and build CSR
    const int64_t kmax = std::max(nm2l, np2p);
    VPMX_CHK(grow(&B.keys_tmp, B.cap_tmpk, std::max<int64_t>(kmax, 1), st));
    size_t need = 0;
    VPMX_CHK(rocprim::radix_sort_keys(nullptr, need, B.m2l_keys, B.keys_tmp, (unsigned int)kmax, nb,
                                      2 * nb, st));
    if (need > B.tmp_bytes) {
        if (B.tmp) (void)hipFree(B.tmp);
        B.tmp = nullptr;
        VPMX_CHK(hipMalloc((void **)&B.tmp, need));
        B.tmp_bytes = need;
    }
    // lists into FmmState (off: ncell + 1, src: count)
    if (ncell + 1 > S.cap_cells + 1 || !S.m2l_off) return hipErrorInvalidValue;
    VPMX_CHK(grow(&S.m2l_src, S.cap_m2l, std::max<int64_t>(nm2l, 1), st));
    VPMX_CHK(grow(&S.p2p_src, S.cap_p2p, std::max<int64_t>(np2p, 1), st));
    struct L { uint64_t *keys; int64_t n; int64_t *off; int32_t *src; } lists[2] = {
        {B.m2l_keys, nm2l, S.m2l_off, S.m2l_src}, {B.p2p_keys, np2p, S.p2p_off, S.p2p_src}};
    for (auto &li : lists) {
        if (li.n > 0) {
            size_t tb = B.tmp_bytes;
            VPMX_CHK(rocprim::radix_sort_keys(B.tmp, tb, li.keys, B.keys_tmp, (unsigned int)li.n, nb,
                                              2 * nb, st));
        }
        const int64_t work = std::max(li.n, ncell + 1);
        csr_kernel<<<gridn(work, FMM_BLOCK), FMM_BLOCK, 0, st>>>(B.keys_tmp, li.n, ncell, nb, li.off,
                                                                li.src);
        VPMX_CHK(hipGetLastError());
    }
    VPMX_CHK(hipMemsetAsync(B.totals + 3, 0, sizeof(int64_t), st));
    // B.keys_tmp holds the sorted P2P keys (the last list above)
    if (np2p > 0)
        p2p_count_kernel<<<(unsigned)std::min<int64_t>(gridn(np2p, FMM_BLOCK), 2048), FMM_BLOCK, 0,
                           st>>>(B.keys_tmp, np2p, nb, S.cells, (unsigned long long *)(B.totals + 3));
    VPMX_CHK(hipGetLastError());
    VPMX_CHK(hipMemcpyAsync(B.h_totals + 3, B.totals + 3, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    VPMX_CHK(hipStreamSynchronize(st));
    res.nm2l = nm2l;
    res.np2p = np2p;
    res.p2p_pairs = B.h_totals[3];
    return hipSuccess;
}

// Near-field work items: a leaf's P2P list in runs of at most P2P_CHUNK
// entries; leaves with several runs get partial-sum slots (their particle count
// per run).  Counts, exclusive scans (offsets and bases, with the totals at
// index nleaf), then the items.
__global__ __launch_bounds__(FMM_BLOCK) void p2p_chunk_count_kernel(
    const int32_t *__restrict__ leaves, int64_t nleaf, const int64_t *__restrict__ off,
    const FmmCellDev *__restrict__ cells, int64_t *__restrict__ nch, int64_t *__restrict__ nslot)
{
    const int64_t li = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (li > nleaf) return;
    if (li == nleaf) {
        nch[li] = 0;
        nslot[li] = 0;
        return;
    }
    const int c = leaves[li];
    const int64_t ne = off[c + 1] - off[c];
    const int64_t n = ne > P2P_CHUNK ? (ne + P2P_CHUNK - 1) / P2P_CHUNK : 1;
    nch[li] = n;
    nslot[li] = n > 1 ? n * (int64_t)(cells[c].end - cells[c].begin) : 0;
}

__global__ __launch_bounds__(FMM_BLOCK) void p2p_chunk_fill_kernel(
    const int32_t *__restrict__ leaves, int64_t nleaf, const int64_t *__restrict__ off,
    const FmmCellDev *__restrict__ cells, const int64_t *__restrict__ coff,
    const int64_t *__restrict__ pbase, P2PChunk *__restrict__ chunks)
{
    const int64_t li = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (li >= nleaf) return;
    const int c = leaves[li];
    const int64_t n = coff[li + 1] - coff[li], nt = cells[c].end - cells[c].begin;
    for (int64_t j = 0; j < n; ++j) {
        P2PChunk q;
        q.c = c;
        q.pad = 0;
        q.e0 = off[c] + j * P2P_CHUNK;
        q.e1 = min(off[c + 1], q.e0 + P2P_CHUNK);
        q.part = n > 1 ? pbase[li] + j * nt : -1;
        chunks[coff[li] + j] = q;
    }
}

// ---------------------------------------------------------------------------
// Device octree
// ---------------------------------------------------------------------------
// The tree of fmm_build_tree, built on the device from the sorted 63-bit keys
// level by level: every cell of a level holding more than ncrit particles
// (and above the deepest level) is cut into its non-empty octants by binary
// searches of the key range; the children of a level are appended in parent
// order, each parent's in octant order (exclusive scans of the child counts),
// so the breadth-first cell order and every cell are those of the host build.
// One read-back per level (the next level's size and its number of parents).
__device__ __forceinline__ int32_t key_lower_bound(const uint64_t *__restrict__ key, int32_t b,
                                                   int32_t e, uint64_t v)
{
    while (b < e) {
        const int32_t m = b + (e - b) / 2;
        if (key[m] < v) b = m + 1; else e = m;
    }
    return b;
}

// The root: every particle, level 0, empty prefix, no parent.
__global__ void tree_root_kernel(int32_t n, int32_t *tb, int32_t *te, int32_t *lev, uint64_t *pre,
                                 int32_t *parent)
{
    tb[0] = 0;
    te[0] = n;
    lev[0] = 0;
    pre[0] = 0;
    parent[0] = -1;
}

// Level cells [lb, lb + m): child count nk[i], split flag, and the ranges and
// octants of the non-empty children; entries nk[m] = flag[m] = 0 close the scans.
__global__ __launch_bounds__(FMM_BLOCK) void tree_split_kernel(
    const uint64_t *__restrict__ key, const int32_t *__restrict__ tb,
    const int32_t *__restrict__ te, const uint64_t *__restrict__ pre,
    const int32_t *__restrict__ lev, int64_t lb, int64_t m, int ncrit, int32_t *__restrict__ nk,
    int32_t *__restrict__ flag, int32_t *__restrict__ kb, int32_t *__restrict__ ke,
    int8_t *__restrict__ ko)
{
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i > m) return;
    if (i == m) {
        nk[m] = 0;
        flag[m] = 0;
        return;
    }
    const int64_t c = lb + i;
    const int32_t b = tb[c], e = te[c], l = lev[c];
    int cnt = 0;
    if (e - b > ncrit && l < KEYBITS) {
        const int shift = 3 * (KEYBITS - l - 1);
        for (uint64_t o = 0; o < 8; o++) {
            const uint64_t cp = (pre[c] << 3) | o;
            const int32_t bb = key_lower_bound(key, b, e, cp << shift);
            const int32_t ee = key_lower_bound(key, bb, e, (cp + 1) << shift);
            if (ee == bb) continue;
            kb[8 * i + cnt] = bb;
            ke[8 * i + cnt] = ee;
            ko[8 * i + cnt] = (int8_t)o;
            cnt++;
        }
    }
    nk[i] = cnt;
    flag[i] = cnt > 0;
}

// Children of the level's cells from slot le + off[i]; parents' child links.
__global__ __launch_bounds__(FMM_BLOCK) void tree_emit_kernel(
    int64_t lb, int64_t m, int64_t le, const int32_t *__restrict__ nk,
    const int32_t *__restrict__ off, const int32_t *__restrict__ kb,
    const int32_t *__restrict__ ke, const int8_t *__restrict__ ko, int32_t *__restrict__ tb,
    int32_t *__restrict__ te, uint64_t *__restrict__ pre, int32_t *__restrict__ lev,
    int32_t *__restrict__ child0, int32_t *__restrict__ nchild, int32_t *__restrict__ parent)
{
    const int64_t i = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (i >= m) return;
    const int64_t c = lb + i;
    const int cnt = nk[i];
    if (cnt == 0) {
        child0[c] = -1;
        nchild[c] = 0;
        return;
    }
    const int64_t first = le + off[i];
    child0[c] = (int32_t)first;
    nchild[c] = cnt;
    for (int r = 0; r < cnt; r++) {
        const int64_t d = first + r;
        tb[d] = kb[8 * i + r];
        te[d] = ke[8 * i + r];
        pre[d] = (pre[c] << 3) | (uint64_t)ko[8 * i + r];
        lev[d] = lev[c] + 1;
        parent[d] = (int32_t)c;
    }
}

// Cell geometry (geometric box centre, half diagonal) and leaves' sigma max;
// leaf / parent flags for the list compaction (entries nc close the scans).
__global__ __launch_bounds__(FMM_BLOCK) void tree_cells_kernel(
    int64_t nc, const int32_t *__restrict__ tb, const int32_t *__restrict__ te,
    const uint64_t *__restrict__ pre, const int32_t *__restrict__ lev,
    const int32_t *__restrict__ child0, const int32_t *__restrict__ nchild,
    const double *__restrict__ ssig, double lo0, double lo1, double lo2, double ext,
    FmmCellDev *__restrict__ cells, int4 *__restrict__ ijk, int32_t *__restrict__ lflag,
    int32_t *__restrict__ pflag)
{
    // product and sum rounded separately, as on the host: the acceptance
    // test compares centre distances exactly
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (c > nc) return;
    if (c == nc) {
        lflag[nc] = 0;
        pflag[nc] = 0;
        return;
    }
    const int l = lev[c];
    const uint64_t p = pre[c];
    uint64_t ix = 0, iy = 0, iz = 0;
    for (int k = 0; k < l; k++) {
        const uint64_t o = (p >> (3 * (l - 1 - k))) & 7u;
        ix = (ix << 1) | (o & 1u);
        iy = (iy << 1) | ((o >> 1) & 1u);
        iz = (iz << 1) | ((o >> 2) & 1u);
    }
    const double h = ext / (double)((uint64_t)1 << l);
    FmmCellDev C;
    C.begin = tb[c];
    C.end = te[c];
    C.child0 = child0[c];
    C.nchild = nchild[c];
    C.cx = lo0 + ((double)ix + 0.5) * h;
    C.cy = lo1 + ((double)iy + 0.5) * h;
    C.cz = lo2 + ((double)iz + 0.5) * h;
    C.R = 0.8660254037844386 * h;
    double s = 0.0;
    if (C.nchild == 0)
        for (int32_t j = C.begin; j < C.end; j++) s = fmax(s, ssig[j]);
    C.smax = s;
    cells[c] = C;
    ijk[c] = make_int4((int)ix, (int)iy, (int)iz, l);
    lflag[c] = C.nchild == 0;
    pflag[c] = C.nchild != 0;
}

// Parents' sigma max from their children, one level at a time from the deepest.
__global__ __launch_bounds__(FMM_BLOCK) void tree_smax_kernel(int64_t lb, int64_t m,
                                                              FmmCellDev *__restrict__ cells)
{
    const int64_t c = lb + (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (c >= lb + m) return;
    const FmmCellDev C = cells[c];
    if (C.nchild == 0) return;
    double s = 0.0;
    for (int k = 0; k < C.nchild; k++) s = fmax(s, cells[C.child0 + k].smax);
    cells[c].smax = s;
}

// leaves and parents in cell order (from the scans of their flags); kids =
// every cell but the root, in cell order.
__global__ __launch_bounds__(FMM_BLOCK) void tree_lists_kernel(
    int64_t nc, const int32_t *__restrict__ lflag, const int32_t *__restrict__ lscan,
    const int32_t *__restrict__ pscan, int32_t *__restrict__ leaves,
    int32_t *__restrict__ parents, int32_t *__restrict__ kids)
{
    const int64_t c = (int64_t)blockIdx.x * FMM_BLOCK + threadIdx.x;
    if (c >= nc) return;
    if (lflag[c])
        leaves[lscan[c]] = (int32_t)c;
    else
        parents[pscan[c]] = (int32_t)c;
    if (c > 0) kids[c - 1] = (int32_t)c;
}

// *p re-allocated for `need` elements, its first `keep` elements copied over.
template <class T>
static hipError_t grow_keep(T **p, int64_t need, int64_t keep, hipStream_t st)
{
    T *q = nullptr;
    hipError_t e = hipMalloc((void **)&q, std::max<int64_t>(need, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    if (*p && keep > 0) {
        e = hipMemcpyAsync(q, *p, keep * sizeof(T), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
    }
    if (*p) (void)hipFree(*p);
    *p = q;
    return hipSuccess;
}

void free_tree_bufs(TreeBufs &B)
{
    for (int32_t **q : {&B.tb, &B.te, &B.lev, &B.child0, &B.nchild, &B.nk, &B.flag, &B.off,
                        &B.foff, &B.kb, &B.ke, &B.lflag, &B.pflag, &B.lscan, &B.pscan})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    if (B.pre) (void)hipFree(B.pre);
    if (B.ko) (void)hipFree(B.ko);
    if (B.scan_tmp) (void)hipFree(B.scan_tmp);
    if (B.h_tot) (void)hipHostFree(B.h_tot);
    B = TreeBufs{};
}

hipError_t fmm_m2l_table(hipStream_t st, FmmState &S, int P)
{
    if (P < 1 || P > M2L_MFMA_MAX_P || S.etab_p == P) return hipSuccess;
    if (!S.m2l_erec) return hipErrorInvalidValue;
    if (S.etab) {
        VPMX_CHK(hipStreamSynchronize(st));
        VPMX_CHK(hipFree(S.etab));
        S.etab = nullptr;
        S.etab_p = -1;
    }
    const int K = m2l_ncoef(P);
    VPMX_CHK(hipMalloc((void **)&S.etab, (size_t)M2L_TAB_ROWS * K * sizeof(double)));
    const unsigned g = gridn(M2L_TAB_ROWS, FMM_BLOCK);
#define VPMX_ETAB(PP)                                                                  \
    case PP: m2l_etab_kernel<PP><<<g, FMM_BLOCK, 0, st>>>(S.m2l_erec, S.etab); break;
    switch (P) {
        VPMX_ETAB(1) VPMX_ETAB(2) VPMX_ETAB(3) VPMX_ETAB(4) VPMX_ETAB(5)
        VPMX_ETAB(6) VPMX_ETAB(7) VPMX_ETAB(8) VPMX_ETAB(9) VPMX_ETAB(10)
    default: return hipErrorInvalidValue;
    }
#undef VPMX_ETAB
    VPMX_CHK(hipGetLastError());
    S.etab_p = P;
    return hipSuccess;
}

hipError_t fmm_build_tree_device(hipStream_t st, FmmState &S, int64_t n, int ncrit,
                                 FmmPlan &pl, int64_t *nc_out)
{
    TreeBufs &B = S.tree;
    if (!B.h_tot) VPMX_CHK(hipHostMalloc((void **)&B.h_tot, 4 * sizeof(int32_t)));
    // cell arrays: grown (keeping the built levels) as levels are added
    auto grow_cells = [&](int64_t need, int64_t keep) -> hipError_t {
        if (need <= B.cap_cell) return hipSuccess;
        const int64_t nc = std::max<int64_t>({need, B.cap_cell + B.cap_cell / 2, 4096});
        VPMX_CHK(grow_keep(&B.tb, nc, keep, st));
        VPMX_CHK(grow_keep(&B.te, nc, keep, st));
        VPMX_CHK(grow_keep(&B.pre, nc, keep, st));
        VPMX_CHK(grow_keep(&B.lev, nc, keep, st));
        VPMX_CHK(grow_keep(&B.child0, nc, keep, st));
        VPMX_CHK(grow_keep(&B.nchild, nc, keep, st));
        VPMX_CHK(grow_keep(&S.parent, nc, keep, st));
        B.cap_cell = nc;
        return hipSuccess;
    };
    auto grow_level = [&](int64_t m) -> hipError_t {
        if (m + 1 <= B.cap_lev) return hipSuccess;
        const int64_t c = std::max<int64_t>({m + 1, B.cap_lev + B.cap_lev / 2, 4096});
        VPMX_CHK(grow_keep(&B.nk, c, 0, st));
        VPMX_CHK(grow_keep(&B.flag, c, 0, st));
        VPMX_CHK(grow_keep(&B.off, c, 0, st));
        VPMX_CHK(grow_keep(&B.foff, c, 0, st));
        VPMX_CHK(grow_keep(&B.kb, 8 * c, 0, st));
        VPMX_CHK(grow_keep(&B.ke, 8 * c, 0, st));
        VPMX_CHK(grow_keep(&B.ko, 8 * c, 0, st));
        B.cap_lev = c;
        return hipSuccess;
    };
    auto scan = [&](const int32_t *in, int32_t *out, int64_t len) -> hipError_t {
        size_t need = 0;
        VPMX_CHK(rocprim::exclusive_scan(nullptr, need, in, out, (int32_t)0, (size_t)len,
                                         rocprim::plus<int32_t>(), st));
        if (need > B.scan_bytes) {
            if (B.scan_tmp) (void)hipFree(B.scan_tmp);
            B.scan_tmp = nullptr;
            B.scan_bytes = 0;
            VPMX_CHK(hipMalloc((void **)&B.scan_tmp, need));
            B.scan_bytes = need;
        }
        size_t bytes = B.scan_bytes;
        return rocprim::exclusive_scan(B.scan_tmp, bytes, in, out, (int32_t)0, (size_t)len,
                                       rocprim::plus<int32_t>(), st);
    };

    VPMX_CHK(grow_cells(1, 0));
    tree_root_kernel<<<1, 1, 0, st>>>((int32_t)n, B.tb, B.te, B.lev, B.pre, S.parent);
    VPMX_CHK(hipGetLastError());
    std::vector<std::pair<int64_t, int64_t>> levels;   // [lb, le) per level
    std::vector<int64_t> nsplit;                       // parents per level
    int64_t lb = 0, le = 1;
    for (;;) {
        const int64_t m = le - lb;
        VPMX_CHK(grow_level(m));
        tree_split_kernel<<<gridn(m + 1, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
            S.kout, B.tb, B.te, B.pre, B.lev, lb, m, ncrit, B.nk, B.flag, B.kb, B.ke, B.ko);
        VPMX_CHK(hipGetLastError());
        VPMX_CHK(scan(B.nk, B.off, m + 1));
        VPMX_CHK(scan(B.flag, B.foff, m + 1));
        VPMX_CHK(hipMemcpyAsync(&B.h_tot[0], B.off + m, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        VPMX_CHK(hipMemcpyAsync(&B.h_tot[1], B.foff + m, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        VPMX_CHK(hipStreamSynchronize(st));
        const int64_t total = B.h_tot[0];
        nsplit.push_back(B.h_tot[1]);
        levels.push_back({lb, le});
        VPMX_CHK(grow_cells(le + total, le));
        tree_emit_kernel<<<gridn(m, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
            lb, m, le, B.nk, B.off, B.kb, B.ke, B.ko, B.tb, B.te, B.pre, B.lev, B.child0, B.nchild,
            S.parent);
        VPMX_CHK(hipGetLastError());
        if (total == 0) break;
        lb = le;
        le += total;
    }
    const int64_t nc = le;
    // device cell arrays and lists
    if (nc > S.cap_cells) {
        const int64_t cap = std::max<int64_t>(nc, S.cap_cells + S.cap_cells / 2);
        VPMX_CHK(grow_keep(&S.cells, cap, 0, st));
        VPMX_CHK(grow_keep(&S.ijk, cap, 0, st));
        VPMX_CHK(grow_keep(&S.leaves, cap, 0, st));
        VPMX_CHK(grow_keep(&S.parents, cap, 0, st));
        VPMX_CHK(grow_keep(&S.kids, cap, 0, st));
        VPMX_CHK(grow_keep(&S.m2l_off, cap + 1, 0, st));
        VPMX_CHK(grow_keep(&S.p2p_off, cap + 1, 0, st));
        S.cap_cells = cap;
    }
    if (nc + 1 > B.cap_flags) {
        const int64_t cap = std::max<int64_t>(nc + 1, B.cap_flags + B.cap_flags / 2);
        VPMX_CHK(grow_keep(&B.lflag, cap, 0, st));
        VPMX_CHK(grow_keep(&B.pflag, cap, 0, st));
        VPMX_CHK(grow_keep(&B.lscan, cap, 0, st));
        VPMX_CHK(grow_keep(&B.pscan, cap, 0, st));
        B.cap_flags = cap;
    }
    int32_t *pflag = B.pflag;
    tree_cells_kernel<<<gridn(nc + 1, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
        nc, B.tb, B.te, B.pre, B.lev, B.child0, B.nchild, S.ssig, pl.lo[0], pl.lo[1], pl.lo[2],
        pl.ext, S.cells, S.ijk, B.lflag, pflag);
    VPMX_CHK(hipGetLastError());
    for (int l = (int)levels.size() - 2; l >= 0; l--) {
        const int64_t m = levels[l].second - levels[l].first;
        tree_smax_kernel<<<gridn(m, FMM_BLOCK), FMM_BLOCK, 0, st>>>(levels[l].first, m, S.cells);
        VPMX_CHK(hipGetLastError());
    }
    VPMX_CHK(scan(B.lflag, B.lscan, nc + 1));
    VPMX_CHK(scan(pflag, B.pscan, nc + 1));
    tree_lists_kernel<<<gridn(nc, FMM_BLOCK), FMM_BLOCK, 0, st>>>(nc, B.lflag, B.lscan, B.pscan,
                                                                  S.leaves, S.parents, S.kids);
    VPMX_CHK(hipGetLastError());
    // host metadata: per-level ranges of parents and kids
    const int nlev = (int)levels.size();
    pl.cells.clear();
    pl.level_parents.assign(nlev, {0, 0});
    pl.level_kids.assign(nlev, {0, 0});
    int64_t np_ = 0;
    for (int l = 0; l < nlev; l++) {
        pl.level_parents[l] = {(int)np_, (int)nsplit[l]};
        np_ += nsplit[l];
        if (l > 0)
            pl.level_kids[l] = {(int)(levels[l].first - 1),
                                (int)(levels[l].second - levels[l].first)};
    }
    pl.nleaves = nc - np_;
    pl.ncells = nc;
    *nc_out = nc;
    return hipSuccess;
}

hipError_t fmm_p2p_chunks(hipStream_t st, FmmState &S, int64_t nleaf)
{
    S.nchunk = S.nslot = 0;
    if (nleaf == 0) return hipSuccess;
    if (nleaf > S.cap_leaf || !S.chunk_buf) {
        const int64_t cap = std::max<int64_t>(nleaf, S.cap_leaf + S.cap_leaf / 2);
        if (S.chunk_buf) (void)hipFree(S.chunk_buf);
        S.chunk_buf = nullptr;
        VPMX_CHK(hipMalloc((void **)&S.chunk_buf, 4 * (cap + 1) * sizeof(int64_t)));
        S.cap_leaf = cap;
    }
    if (!S.h_chunk_tot)
        VPMX_CHK(hipHostMalloc((void **)&S.h_chunk_tot, 2 * sizeof(int64_t), hipHostMallocDefault));
    const int64_t L1 = S.cap_leaf + 1;
    int64_t *ncnt = S.chunk_buf, *scnt = S.chunk_buf + 2 * L1;
    S.chunk_off = S.chunk_buf + L1;
    S.part_base = S.chunk_buf + 3 * L1;
    p2p_chunk_count_kernel<<<gridn(nleaf + 1, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
        S.leaves, nleaf, S.p2p_off, S.cells, ncnt, scnt);
    VPMX_CHK(hipGetLastError());
    size_t need = 0;
    VPMX_CHK(rocprim::exclusive_scan(nullptr, need, ncnt, S.chunk_off, (int64_t)0,
                                     (size_t)(nleaf + 1), rocprim::plus<int64_t>(), st));
    if (need > S.chunk_tmp_bytes) {
        if (S.chunk_tmp) (void)hipFree(S.chunk_tmp);
        S.chunk_tmp = nullptr;
        VPMX_CHK(hipMalloc((void **)&S.chunk_tmp, need));
        S.chunk_tmp_bytes = need;
    }
    size_t tb = S.chunk_tmp_bytes;
    VPMX_CHK(rocprim::exclusive_scan(S.chunk_tmp, tb, ncnt, S.chunk_off, (int64_t)0,
                                     (size_t)(nleaf + 1), rocprim::plus<int64_t>(), st));
    tb = S.chunk_tmp_bytes;
    VPMX_CHK(rocprim::exclusive_scan(S.chunk_tmp, tb, scnt, S.part_base, (int64_t)0,
                                     (size_t)(nleaf + 1), rocprim::plus<int64_t>(), st));
    VPMX_CHK(hipMemcpyAsync(S.h_chunk_tot, S.chunk_off + nleaf, sizeof(int64_t),
                            hipMemcpyDeviceToHost, st));
    VPMX_CHK(hipMemcpyAsync(S.h_chunk_tot + 1, S.part_base + nleaf, sizeof(int64_t),
                            hipMemcpyDeviceToHost, st));
    VPMX_CHK(hipStreamSynchronize(st));
    S.nchunk = S.h_chunk_tot[0];
    S.nslot = S.h_chunk_tot[1];
    VPMX_CHK(grow(&S.chunks, S.cap_chunk, std::max<int64_t>(S.nchunk, 1), st));
    VPMX_CHK(grow(&S.part, S.cap_part, std::max<int64_t>(14 * S.nslot, 1), st));
    p2p_chunk_fill_kernel<<<gridn(nleaf, FMM_BLOCK), FMM_BLOCK, 0, st>>>(
        S.leaves, nleaf, S.p2p_off, S.cells, S.chunk_off, S.part_base, S.chunks);
    return hipGetLastError();
}

#undef VPMX_CHK

// Host version of the key computation and the whole plan, for checking the
// tree and traversal against an independent CPU implementation.
static uint64_t spread21_host(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

void fmm_plan_host(FmmPlan &pl, int64_t n, const double *X, const double *sig, int ncrit,
                   double theta, double sigcut)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], X[3 * i + k]);
            hi[k] = fmax(hi[k], X[3 * i + k]);
        }
    double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    if (!(ext > 0.0)) ext = 1.0;
    ext *= 1.0000001;
    const double scale = 2097152.0 / ext;
    std::vector<std::pair<uint64_t, int64_t>> ki(n);
    for (int64_t i = 0; i < n; i++) {
        uint64_t q[3];
        for (int k = 0; k < 3; k++) {
            double v = (X[3 * i + k] - lo[k]) * scale;
            v = fmin(fmax(v, 0.0), 2097151.0);
            q[k] = (uint64_t)v;
        }
        ki[i] = {spread21_host(q[0]) | (spread21_host(q[1]) << 1) | (spread21_host(q[2]) << 2), i};
    }
    std::sort(ki.begin(), ki.end());
    std::vector<uint64_t> key(n);
    std::vector<double> ss(n);
    for (int64_t i = 0; i < n; i++) {
        key[i] = ki[i].first;
        ss[i] = sig[ki[i].second];
    }
    for (int k = 0; k < 3; k++) pl.lo[k] = lo[k];
    pl.ext = ext;
    fmm_build_tree(pl, key.data(), ss.data(), n, ncrit);
    fmm_traverse(pl, theta, sigcut);
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
hipError_t launch_fmm_keys(hipStream_t st, const DeviceField &f, const double lo[3], double scale,
                           uint64_t *kin, int32_t *vin)
{
    if (f.np == 0) return hipSuccess;
    morton63_kernel<<<gridn(f.np, FMM_BLOCK), FMM_BLOCK, 0, st>>>(f.X, f.np, f.cap, lo[0], lo[1],
                                                                   lo[2], scale, kin, vin);
    return hipGetLastError();
}

size_t fmm_sort_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, bytes, (const uint64_t *)nullptr,
                                    (uint64_t *)nullptr, (const int32_t *)nullptr,
                                    (int32_t *)nullptr, (unsigned int)n, 0, 63, (hipStream_t)0);
    return bytes;
}

hipError_t launch_fmm_sort(hipStream_t st, void *tmp, size_t bytes, const uint64_t *kin,
                           uint64_t *kout, const int32_t *vin, int32_t *perm, int64_t n)
{
    if (n == 0) return hipSuccess;
    return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, perm, (unsigned int)n, 0, 63, st);
}

hipError_t launch_fmm_gather(hipStream_t st, const DeviceField &f, const int32_t *perm,
                             double4 *rec, double *ssig)
{
    if (f.np == 0) return hipSuccess;
    fmm_gather_kernel<<<gridn(f.np, FMM_BLOCK), FMM_BLOCK, 0, st>>>(f.X, f.G, f.sigma, perm, f.np,
                                                                     f.cap, rec, ssig);
    return hipGetLastError();
}

// Grid of the persistent matrix-core M2L: resident blocks per CU (from the
// kernel's register and LDS use) times the CUs, a multiple of 8 (one share
// per XCD), at most one block per 8 cells.
static int m2l_mfma_grid(const void *fn, int P, int ncell)
{
    static int cached[M2L_MFMA_MAX_P + 1] = {0};
    if (!cached[P]) {
        int dev = 0, ncu = 0, nb = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64, 0) != hipSuccess)
            ncu = 256, nb = 4;
        cached[P] = std::max(1, ncu) * std::max(1, nb);
    }
    const int want = std::min(cached[P], std::max(1, ncell / 8));
    return (want + 7) / 8 * 8;
}

hipError_t launch_fmm_eval(hipStream_t st, const FmmDev &d, const FmmPlan &pl, int kernel,
                           int precision, const DeviceField &f, int reset)
{
    const int K = d.K, P = d.P;
    const size_t shK = sizeof(double) * 4 * K;
    if (d.nleaf > 0) {
        p2m_kernel<<<d.nleaf, FMM_BLOCK, 0, st>>>(d.leaves, d.cells, d.rec, K, d.mi, d.M);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // M2M bottom-up: parents level by level
    for (int lev = (int)pl.level_parents.size() - 1; lev >= 0; lev--) {
        const int cnt = pl.level_parents[lev].second;
        if (cnt == 0) continue;
        m2m_kernel<<<cnt, FMM_BLOCK, shK, st>>>(d.parents + pl.level_parents[lev].first, d.cells, K,
                                               d.mi, d.m2m_off, d.m2m_idx, d.m2m_coef, d.M);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (d.ncell > 0 && P >= 1 && P <= M2L_MFMA_MAX_P) {
        const int64_t nm = (int64_t)d.ncell * 3 * K;
        if (!d.ijk || !d.etab) return hipErrorInvalidValue;
        m2l_prescale_kernel<<<gridn(nm, FMM_BLOCK), FMM_BLOCK, 0, st>>>(d.M, nm, K, d.m2l_bscale,
                                                                       d.ijk, d.m2l_erec, d.ext);
        // the zero multipole row read for absent sources (M holds ncell + 1 rows)
        fmm_zero_kernel<<<gridn(3 * K, FMM_BLOCK), FMM_BLOCK, 0, st>>>(d.M + nm, 3 * K);
#define VPMX_M2LM(PP)                                                                           \
    case PP: {                                                                                  \
        const int grid = m2l_mfma_grid((const void *)m2l_mfma_kernel<PP>, PP, d.ncell);          \
        m2l_mfma_kernel<PP><<<grid, 64, 0, st>>>(d.ijk, d.ncell, d.m2l_off, d.m2l_src,           \
                                                 d.m2l_erec, d.m2l_tile_idx, d.m2l_lscale,      \
                                                 d.etab, d.ext, d.M, (int64_t)d.ncell, d.L);    \
        break;                                                                                  \
    }
        switch (P) {
            VPMX_M2LM(1) VPMX_M2LM(2) VPMX_M2LM(3) VPMX_M2LM(4) VPMX_M2LM(5)
            VPMX_M2LM(6) VPMX_M2LM(7) VPMX_M2LM(8) VPMX_M2LM(9) VPMX_M2LM(10)
        default: return hipErrorInvalidValue;
        }
#undef VPMX_M2LM
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    } else if (d.ncell > 0) {
        // derivatives + multipoles of a batch, or the chunk sums at the end
        const size_t shm = sizeof(double) * std::max<size_t>((size_t)M2L_BATCH * 4 * K,
                                                             (size_t)3 * d.m2l_nchunk);
#define VPMX_M2L(RR)                                                                           \
    m2l_kernel<RR><<<d.ncell, FMM_BLOCK, shm, st>>>(d.cells, d.m2l_off, d.m2l_src, K, P, d.rec1, \
                                                    d.rec2, d.m2l_chunks, d.m2l_nchunk,         \
                                                    d.m2l_cfirst, d.m2l_tidx, d.m2l_tcoef, d.M, d.L)
        switch (d.m2l_rounds) {
        case 1: VPMX_M2L(1); break;
        case 2: VPMX_M2L(2); break;
        case 4: VPMX_M2L(4); break;
        case 8: VPMX_M2L(8); break;
        default: return hipErrorInvalidValue;
        }
#undef VPMX_M2L
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (size_t lev = 1; lev < pl.level_kids.size(); lev++) {
        const int cnt = pl.level_kids[lev].second;
        if (cnt == 0) continue;
        l2l_kernel<<<cnt, FMM_BLOCK, shK, st>>>(d.kids + pl.level_kids[lev].first, d.parent, d.cells,
                                               K, d.mi, d.l2l_off, d.l2l_idx, d.l2l_coef, d.L);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (d.nleaf > 0) {
        const size_t sh = sizeof(double) * (3 * K + P2P_BLOCK * 3 * (P + 1));
        l2p_kernel<<<d.nleaf, P2P_BLOCK, sh, st>>>(d.leaves, d.cells, d.rec, d.perm, K, P, d.mi,
                                                  d.d1, d.d2a, d.d2b, d.L, f.cap, f.U, f.J, reset);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const unsigned g = (unsigned)d.nchunk;
#define VPMX_P2P(KK)                                                                              \
    if (precision == 0)                                                                           \
        p2p_f64_kernel<KK><<<g, P2P_BLOCK, 0, st>>>(d.chunks, d.cells, d.p2p_src, d.rec, d.perm,   \
                                                   f.cap, f.U, f.J, d.part, d.nslot);             \
    else                                                                                          \
        p2p_mixed_kernel<KK><<<g, P2P_BLOCK, 0, st>>>(d.chunks, d.cells, d.p2p_src, d.rec, d.perm, \
                                                     f.cap, f.U, f.J, d.part, d.nslot);
        switch (kernel) {
        case SINGULAR: VPMX_P2P(SINGULAR); break;
        case GAUSSIAN: VPMX_P2P(GAUSSIAN); break;
        case GAUSSIANERF: VPMX_P2P(GAUSSIANERF); break;
        case WINCKELMANS: VPMX_P2P(WINCKELMANS); break;
        default: return hipErrorInvalidValue;
        }
#undef VPMX_P2P
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = precision == 0
                ? launch_part_reduce<0>(st, d.leaves, d.nleaf, d.chunk_off, d.part_base, d.cells,
                                        d.part, d.nslot, d.perm, f.cap, f.U, f.J, nullptr)
                : launch_part_reduce<1>(st, d.leaves, d.nleaf, d.chunk_off, d.part_base, d.cells,
                                        d.part, d.nslot, d.perm, f.cap, f.U, f.J, nullptr);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace vpmx
