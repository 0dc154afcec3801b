// fmm.h -- tree, traversal lists, operator tables and device views of the
// fast multipole UJ (fmm.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "internal.h"

namespace vpmx {

constexpr int FMM_MAX_P = 12;   // K = 455 <= 2 coefficients per thread of 256

// M2L on the matrix cores (orders 1..M2L_MFMA_MAX_P): the operator of one
// source, T[b][a] = E_{a+b} on the region |a| + |b| <= P, cut into 4 x 4 tiles
// (rows b = 4 tb + 0..3, columns a = 4 ta + 0..3).  Coefficients are in graded
// order, so the row length of b -- the number of a with |a| <= P - |b| -- only
// falls with b: b-tile tb needs a-tiles 0 .. m2l_natile(P, tb) - 1, and a-tile
// ta meets b-tiles 0 .. m2l_nrows(P, ta) - 1.  Tiles are numbered a-tile by
// a-tile: tile (tb, ta) is m2l_base(P, ta) + tb.
constexpr int M2L_MFMA_MAX_P = 10;
constexpr int m2l_ncoef(int p) { return p < 0 ? 0 : (p + 1) * (p + 2) * (p + 3) / 6; }
constexpr int m2l_degree(int k)
{
    int n = 0;
    while (m2l_ncoef(n) <= k) ++n;
    return n;
}
constexpr int m2l_ntile1(int P) { return (m2l_ncoef(P) + 3) / 4; }
constexpr int m2l_natile(int P, int tb)
{
    return 4 * tb < m2l_ncoef(P) ? (m2l_ncoef(P - m2l_degree(4 * tb)) + 3) / 4 : 0;
}
constexpr int m2l_nrows(int P, int ta)
{
    int n = 0;
    for (int tb = 0; tb < m2l_ntile1(P); ++tb)
        if (m2l_natile(P, tb) > ta) ++n;
    return n;
}
constexpr int m2l_base(int P, int ta)
{
    int s = 0;
    for (int t = 0; t < ta; ++t) s += m2l_nrows(P, t);
    return s;
}
// Total tiles: sum over b-tiles of their a-tile counts (= sum of m2l_nrows).
constexpr int m2l_ntiles(int P)
{
    int s = 0;
    for (int tb = 0; tb < m2l_ntile1(P); ++tb) s += m2l_natile(P, tb);
    return s;
}
// Tile count padded to whole batches of 8 (the index table's row length).
constexpr int m2l_ntiles8(int P) { return (m2l_ntiles(P) + 7) / 8 * 8; }

struct FmmCellDev {
    int32_t begin, end, child0, nchild;
    double cx, cy, cz, R, smax;
};

// Operator table of the matrix-core M2L.  A cell of level l has centre
// lo + (i + 1/2) ext 2^-l (integer coordinates i, kept per cell as (ix, iy,
// iz, l)), so a pair's separation is (ext 2^-(lf+1)) v with v an integer
// vector in half-widths of the finer cell (level lf), and its derivative
// operator is a scaled copy of E(v).  Rows are kept for |v_i| <= M2L_TAB_V
// and target-side level gaps dA = lf - l_target <= M2L_TAB_D; other pairs
// derive E by the recurrence.
constexpr int M2L_TAB_V = 23, M2L_TAB_D = 2, M2L_TAB_W = 2 * M2L_TAB_V + 1;
constexpr int64_t M2L_TAB_ROWS = (int64_t)(M2L_TAB_D + 1) * M2L_TAB_W * M2L_TAB_W * M2L_TAB_W;

// Near-field work item: entries [e0, e1) of target leaf c's P2P list.  A leaf
// whose list is longer than P2P_CHUNK entries is cut into several items (one
// workgroup each), which write FP64 partial sums from slot `part` (the leaf's
// particles, one slot each) for a fixed-order reduction; part < 0: the item is
// the leaf's whole list and stores directly.
struct P2PChunk {
    int64_t e0, e1, part;
    int32_t c, pad;
};
constexpr int P2P_CHUNK = 512;

// Scratch of the device dual-tree traversal.
struct DttBufs {
    int64_t cap_front = 0, cap_next = 0, cap_m2l = 0, cap_p2p = 0, cap_count = 0, cap_tmpk = 0;
    int2 *front = nullptr, *next = nullptr;
    int *cnt_next = nullptr, *cnt_m2l = nullptr, *cnt_p2p = nullptr;
    int64_t *off_next = nullptr, *off_m2l = nullptr, *off_p2p = nullptr;
    uint64_t *m2l_keys = nullptr, *p2p_keys = nullptr, *keys_tmp = nullptr;
    int64_t *totals = nullptr;               // 3 counts + pair total
    int64_t *h_totals = nullptr;             // pinned
    uint8_t *tmp = nullptr;
    size_t tmp_bytes = 0;
};

// Scratch of the device octree build: per-cell level arrays (grown as levels
// are added), per-level split counts, child ranges and octants, the scans of
// the leaf / parent flags, and a pinned pair of level totals.
struct TreeBufs {
    int64_t cap_cell = 0, cap_lev = 0, cap_flags = 0;
    int32_t *tb = nullptr, *te = nullptr, *lev = nullptr, *child0 = nullptr, *nchild = nullptr;
    uint64_t *pre = nullptr;
    int32_t *nk = nullptr, *flag = nullptr, *off = nullptr, *foff = nullptr;
    int32_t *kb = nullptr, *ke = nullptr;
    int8_t *ko = nullptr;
    int32_t *lflag = nullptr, *pflag = nullptr, *lscan = nullptr, *pscan = nullptr;
    uint8_t *scan_tmp = nullptr;
    size_t scan_bytes = 0;
    int32_t *h_tot = nullptr;
};

// Result sizes of a device traversal.
struct DttResult {
    int64_t nm2l = 0, np2p = 0, p2p_pairs = 0;
    int steps = 0;
};

struct FmmCellHost {
    int64_t begin, end, child0;
    int nchild, level;
    uint64_t prefix;
    double c[3], R, smax;
};

// Operator tables of expansion order P (host copies).
struct FmmTables {
    int P = -1, K = 0;
    std::vector<int> mi, rec1, rec2, d1, d2a, d2b;
    std::vector<int> m2m_off, m2m_idx, l2l_off, l2l_idx, m2l_off, m2l_idx;
    std::vector<double> m2m_coef, l2l_coef, m2l_coef;
    // M2L work split: chunks (b, first term, end term, 0) of at most m2l_chunk_len
    // terms of one output coefficient b, in b order; m2l_cfirst[b] = first chunk
    // of b; m2l_rounds = chunks per thread of a 256-thread block (1, 2, 4 or 8).
    std::vector<int> m2l_chunk, m2l_cfirst;
    int m2l_nchunk = 0, m2l_rounds = 1, m2l_chunk_len = 0;
    // Matrix-core M2L (P <= M2L_MFMA_MAX_P): for element (row, k) of tile t, at
    // (row + 4 k) m2l_ntiles8(P) + t, the index of a + b in E, or K (a zero
    // slot) outside the region and in the padding; bscale[a] = (-1)^|a| / a!,
    // lscale[b] = 1 / b!.
    std::vector<uint16_t> m2l_tile_idx;
    std::vector<double> m2l_bscale, m2l_lscale;
    std::vector<int> m2l_erec;   // 4 per k, packed as m2l_mfma_kernel reads them
};

// Tree + lists of one evaluation (host).
struct FmmPlan {
    double lo[3] = {0, 0, 0}, ext = 1.0;
    std::vector<FmmCellHost> cells;
    std::vector<int64_t> m2l_off, p2p_off;
    std::vector<int32_t> m2l_src, p2p_src;
    std::vector<int32_t> leaves, parents, kids, parent;
    std::vector<std::pair<int, int>> level_parents, level_kids;   // (offset, count) per level
    int64_t p2p_pairs = 0;
    int64_t ncells = 0, nleaves = 0;   // of the device-built tree
    double ms_tree = 0, ms_traverse = 0, ms_upload = 0;
};

// Device views used by the kernels.
struct FmmDev {
    int P = 0, K = 0, ncell = 0, nleaf = 0;
    const FmmCellDev *cells = nullptr;
    const int32_t *leaves = nullptr, *parents = nullptr, *kids = nullptr, *parent = nullptr;
    const int64_t *m2l_off = nullptr, *p2p_off = nullptr;
    const int32_t *m2l_src = nullptr, *p2p_src = nullptr;
    const double4 *rec = nullptr;      // sorted source records
    const int32_t *perm = nullptr;     // sorted slot -> original index
    double *M = nullptr, *L = nullptr;
    const int *mi = nullptr;
    const int4 *rec1 = nullptr, *rec2 = nullptr, *d1 = nullptr, *d2a = nullptr, *d2b = nullptr;
    const int *m2m_off = nullptr, *l2l_off = nullptr, *m2l_toff = nullptr;
    const int2 *m2m_idx = nullptr, *l2l_idx = nullptr, *m2l_tidx = nullptr;
    const double *m2m_coef = nullptr, *l2l_coef = nullptr, *m2l_tcoef = nullptr;
    const int4 *m2l_chunks = nullptr;
    const int *m2l_cfirst = nullptr;
    int m2l_nchunk = 0, m2l_rounds = 1;
    const uint16_t *m2l_tile_idx = nullptr;
    const double *m2l_bscale = nullptr, *m2l_lscale = nullptr;
    const int4 *m2l_erec = nullptr;
    const int4 *ijk = nullptr;        // per cell (ix, iy, iz, level)
    const double *etab = nullptr;     // M2L_TAB_ROWS x K operator rows
    double ext = 1.0;                 // edge of the root cell
    // near-field work items (fmm_p2p_chunks)
    const P2PChunk *chunks = nullptr;
    const int64_t *chunk_off = nullptr, *part_base = nullptr;
    double *part = nullptr;
    int64_t nchunk = 0, nslot = 0;
};

// Per-context FMM state: host plan, operator tables and device buffers.
struct FmmState {
    FmmPlan plan;
    FmmTables tab;
    int dev_tab_p = -1;
    // device operator tables
    int *mi = nullptr, *m2m_off = nullptr, *l2l_off = nullptr, *m2l_toff = nullptr;
    int4 *rec1 = nullptr, *rec2 = nullptr, *d1 = nullptr, *d2a = nullptr, *d2b = nullptr;
    int2 *m2m_idx = nullptr, *l2l_idx = nullptr, *m2l_tidx = nullptr;
    double *m2m_coef = nullptr, *l2l_coef = nullptr, *m2l_tcoef = nullptr;
    int4 *m2l_chunks = nullptr;
    int *m2l_cfirst = nullptr;
    uint16_t *m2l_tile_idx = nullptr;
    double *m2l_bscale = nullptr, *m2l_lscale = nullptr;
    int4 *m2l_erec = nullptr;
    // near-field work items: per leaf chunk count / partial slots (scanned in
    // place into offsets, nleaf + 1 each), the items, the partial sums
    int64_t cap_leaf = 0, cap_chunk = 0, cap_part = 0, nchunk = 0, nslot = 0;
    int64_t *chunk_buf = nullptr;   // 4 (cap_leaf + 1): counts, offsets, slot counts, bases
    int64_t *chunk_off = nullptr, *part_base = nullptr, *h_chunk_tot = nullptr;
    P2PChunk *chunks = nullptr;
    double *part = nullptr;
    uint8_t *chunk_tmp = nullptr;
    size_t chunk_tmp_bytes = 0;
    // device tree and lists
    int64_t cap_cells = 0, cap_m2l = 0, cap_p2p = 0, cap_coef = 0, cap_n = 0;
    FmmCellDev *cells = nullptr;
    int4 *ijk = nullptr;
    int32_t *leaves = nullptr, *parents = nullptr, *kids = nullptr, *parent = nullptr;
    int64_t *m2l_off = nullptr, *p2p_off = nullptr;
    int32_t *m2l_src = nullptr, *p2p_src = nullptr;
    double *M = nullptr, *L = nullptr;
    // sorting
    uint64_t *kin = nullptr, *kout = nullptr;
    double *ssig = nullptr;
    uint8_t *sort_tmp = nullptr;
    size_t sort_bytes = 0;
    // matrix-core M2L operator rows of order etab_p
    double *etab = nullptr;
    int etab_p = -1;
    // statistics of the last evaluation
    double stats[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    DttBufs dtt;
    TreeBufs tree;
};

// Estr over the near-field lists of the last FMM evaluation (J must be current);
// precision 1 runs the FP32 pair math with FP64 accumulation.  With DJ
// (9 x cap, the core-size derivative of J) also D[E] into DE.
hipError_t launch_fmm_estr(hipStream_t st, const FmmState &S, int kernel, int transposed,
                           int precision, const DeviceField &f, Scratch &s,
                           const double *DJ = nullptr, double *DE = nullptr);
// Viscous sums over the near-field lists of the last FMM evaluation (FP64,
// outputs overwritten, natural particle order): the vorticity
// w = sum_j G_j zeta_s(x - x_j) (s = sigma_j, or s0 > 0 for every source;
// sources in `mask` without strength; G and w of strides gld and wld) and
// the PSE rate into f.V.  `rec` holds 2 records per particle.
hipError_t launch_fmm_zeta(hipStream_t st, const FmmState &S, int kernel, const DeviceField &f,
                           const Scratch &s, const double *G, int64_t gld, double s0,
                           const uint8_t *mask, double4 *rec, double *w, int64_t wld);
hipError_t launch_fmm_pse(hipStream_t st, const FmmState &S, int kernel, const DeviceField &f,
                          const Scratch &s, double nu, double4 *rec);
// Core-size derivatives D[U] (3 x cap) and D[J] (9 x cap) over the near-field
// lists of the last FMM evaluation (overwritten; the far field has none).
hipError_t launch_fmm_dsig(hipStream_t st, const FmmState &S, int kernel, int precision,
                           const DeviceField &f, const Scratch &s, double *DU, double *DJ);

// Near-field work items of the current lists (S.chunks, S.nchunk, S.nslot;
// S.part sized for 14 values per slot).  nleaf = S.plan.nleaves.
hipError_t fmm_p2p_chunks(hipStream_t st, FmmState &S, int64_t nleaf);

// Dual tree traversal on the device: fills m2l_off/m2l_src and p2p_off/p2p_src
// (CSR by target cell, each target's sources in emission order), growing the
// buffers in S.
hipError_t fmm_traverse_device(hipStream_t st, FmmState &S, int64_t ncell, double theta,
                               double sigcut, DttResult &res);

// Octree of the sorted keys S.kout / sigmas S.ssig (n particles, bounding cube
// pl.lo, pl.ext) built on the device: S.cells, S.leaves, S.parents, S.kids and
// S.parent (grown as needed) and, on the host, pl.level_parents / level_kids,
// pl.ncells and pl.nleaves.  The same tree as fmm_build_tree.
hipError_t fmm_build_tree_device(hipStream_t st, FmmState &S, int64_t n, int ncrit, FmmPlan &pl,
                                 int64_t *nc_out);
void free_tree_bufs(TreeBufs &B);
// The operator table of the matrix-core M2L for order P (1..M2L_MFMA_MAX_P),
// built once per order from the uploaded recurrence records S.m2l_erec.
hipError_t fmm_m2l_table(hipStream_t st, FmmState &S, int P);

void fmm_build_tree(FmmPlan &pl, const uint64_t *key, const double *ssig, int64_t n, int ncrit);
void fmm_traverse(FmmPlan &pl, double theta, double sigcut);
void fmm_build_tables(FmmTables &T, int P);
void fmm_plan_host(FmmPlan &pl, int64_t n, const double *X, const double *sig, int ncrit,
                   double theta, double sigcut);

hipError_t launch_fmm_keys(hipStream_t st, const DeviceField &f, const double lo[3], double scale,
                           uint64_t *kin, int32_t *vin);
size_t fmm_sort_temp_bytes(int64_t n);
hipError_t launch_fmm_sort(hipStream_t st, void *tmp, size_t bytes, const uint64_t *kin,
                           uint64_t *kout, const int32_t *vin, int32_t *perm, int64_t n);
hipError_t launch_fmm_gather(hipStream_t st, const DeviceField &f, const int32_t *perm,
                             double4 *rec, double *ssig);
hipError_t launch_fmm_eval(hipStream_t st, const FmmDev &d, const FmmPlan &pl, int kernel,
                           int precision, const DeviceField &f, int reset);

}  // namespace vpmx
