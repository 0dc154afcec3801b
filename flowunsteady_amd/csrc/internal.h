// internal.h -- context and launcher declarations shared by the library's
// translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "vpmx.h"

namespace vpmx {

// Timed kernel classes (vpmx_get_timing's `which`).
enum TimedKind { T_UJ = 0, T_ESTR = 1, T_UPDATE = 2, T_FMM = 3, T_GRAPH = 4, T_DSIG = 5, T_NKIND = 6 };

struct Timer {
    bool on = false;
    double total_ms[T_NKIND] = {0, 0, 0, 0, 0, 0};
    int64_t count[T_NKIND] = {0, 0, 0, 0, 0, 0};
    struct Pending { int kind; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
};

// Device state of one particle field, SoA FP64 with stride `cap`:
// component k of particle i of a vector field F lives at F[k*cap + i].
struct DeviceField {
    int64_t np = 0, cap = 0;
    double *X = nullptr, *G = nullptr, *sigma = nullptr, *U = nullptr, *J = nullptr;
    double *E = nullptr, *C = nullptr, *M = nullptr;   // C: 3 components
    // dynamic SFS: core-size derivatives (D = sigma d/dsigma of every core)
    // D[U] (3), D[J] (9), D[E] (3)
    double *D = nullptr;
    double *vol = nullptr;     // particle volume (sigma^3 when the layout has none)
    double *V = nullptr;       // viscous: PSE rate of Gamma (3), allocated on first use
    uint8_t *is_static = nullptr;
    bool any_static = false;
    double *DU() const { return D; }
    double *DJ() const { return D + 3 * cap; }
    double *DE() const { return D + 12 * cap; }
};

// Scratch of the pairwise passes, grown with the field.
struct Scratch {
    int64_t cap = 0;
    double4 *src64 = nullptr;        // FP64 UJ source records, 2 per particle
    double4 *srcE = nullptr;         // FP64 Estr source records, 7 per particle
    double4 *spos = nullptr;         // Morton-sorted (x, y, z, p0) FP64
    float4 *sgam = nullptr;          // Morton-sorted (Gx, Gy, Gz, p1) FP32
    double4 *tbnd = nullptr;         // per 256-slot tile of the sorted field:
                                     // (centre, radius), (largest sigma, -, -, -)
    // split source ranges of small launches: partial sums (up to 15 per slot
    // per segment; kcap doubles)
    int64_t kcap = 0;
    double *kpart = nullptr;
    uint32_t *kin = nullptr, *kout = nullptr;
    int32_t *vin = nullptr, *perm = nullptr;
    uint8_t *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    double *part = nullptr;          // per-block partials (bbox, reductions)
    int64_t part_cap = 0;
    // target-subset evaluation (sources = the resident field)
    int64_t tcap = 0;
    double *tX = nullptr, *tU = nullptr, *tJ = nullptr;   // SoA, stride = number of targets
    double4 *tpos = nullptr;         // Morton-sorted targets (x, y, z, 0)
    int32_t *tperm = nullptr;        // sorted slot -> target index
    // viscous schemes: source records, core-spreading CG vectors (3 x vcap
    // each), reduction partials and device scalars
    int64_t vcap = 0;
    double4 *vrec = nullptr;
    double *cw = nullptr, *cr = nullptr, *cp = nullptr, *cAp = nullptr;
    double *cpart = nullptr, *cdev = nullptr;
    // free-vortex force: bound vortex records and per-segment partials
    int64_t bcap = 0, fpcap = 0;
    double *bnd = nullptr, *fpart = nullptr, *fout = nullptr;
};

// One workgroup's work in the batched UJ: targets [t0, t1) against sources
// [s0, s1), all in slots of the concatenated fields.
// Targets [t0, t1) of one field against its sources [s0, s1); the field's
// BLK-slot tiles (from s0) have bounds from tile `tb` of the batch's list.
struct BatchItem { int32_t t0, t1, s0, s1, tb, pad; };

// Schedules of the batched UJ.
enum BatchSchedule { SCHED_FLAT = 0, SCHED_QUEUE = 1 };

// Many independent fields resident as one concatenation: field k owns slots
// [off[k], off[k+1]) of the DeviceField and of every scratch array.
struct BatchState {
    int32_t nfields = 0;             // 0: the resident state is a single field
    std::vector<int64_t> off;        // nfields + 1
    int64_t fcap = 0;                // field capacity of box / uinf / doff
    int64_t *doff = nullptr;         // device copy of off
    double *box = nullptr;           // per field bounding box (lo xyz, hi xyz)
    double *uinf = nullptr;          // per field freestream (3)
    int64_t pcap = 0;                // particle capacity of fid / keys
    int32_t *fid = nullptr;          // per particle field index
    uint64_t *kin = nullptr, *kout = nullptr;   // 64-bit sort keys
    uint8_t *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    // schedule of the pairwise pass and its work list
    int sched = SCHED_QUEUE, block = 256, grid = 0;   // grid 0: resident occupancy x CUs
    int grid_eff = 0;                // grid of the last queue launch setup
    bool tuned = false;
    int items_block = 0;             // block size the work list was built for
    std::vector<int64_t> items_off;  // offsets the work list was built for
    int64_t icap = 0, nitems = 0;
    BatchItem *items = nullptr;
    int32_t *counter = nullptr;      // work-queue head
    // per-field tiles of items_block slots (slot ranges) and their bounds
    int64_t tcap = 0, ntiles = 0;
    int2 *tiles = nullptr;
    double4 *tbnd = nullptr;
};

// Instantiated hipGraph of one time step on the resident state and the
// settings it was captured with.
struct GraphState {
    bool on = false;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
};

struct FmmState;   // fmm.h

}  // namespace vpmx

struct vpmx_ctx {
    vpmx::FmmState *fmm = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int inject = 0;
    vpmx::DeviceField f;
    vpmx::Scratch s;
    vpmx::Timer timer;
    vpmx::BatchState b;
    vpmx::GraphState g;
    bool capturing = false;          // enqueueing into a graph: no events, no host syncs
    uint64_t gen = 0;                // bumped whenever a device buffer is reallocated
    int ncu = 0;                     // compute units of the device
    double *h_stage = nullptr;       // pinned host staging
    size_t h_stage_bytes = 0;
    double *h_small = nullptr;       // pinned staging of per-step scalars (freestreams)
    size_t h_small_bytes = 0;
};

namespace vpmx {

// Launchers (kernels.hip).  All enqueue on `st` and return hipSuccess or the
// launch error.
hipError_t launch_pack_f64(hipStream_t st, const DeviceField &f, Scratch &s);
hipError_t launch_uj_f64(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                         int reset);
// FP64 direct UJ on caller-owned SoA device arrays of stride n, with its
// scratch (source records, split partials) in a caller-provided block of at
// least uj_f64_dev_scratch_bytes(n) bytes, 256-byte aligned.
size_t uj_f64_dev_scratch_bytes(int64_t n);
hipError_t launch_uj_f64_dev(hipStream_t st, int kernel, int64_t n, const double *X,
                             const double *G, const double *sigma, double *U, double *J,
                             int reset, void *scratch);
// FP64 direct UJ of a single field on its Morton order (launch_sort_mixed's
// perm): sorted source records into s.src64 and tile bounds into s.tbnd, then
// the pair kernel with the far-field branch, results stored through perm.
hipError_t launch_pack_f64_sorted(hipStream_t st, const DeviceField &f, Scratch &s);
hipError_t launch_uj_f64_sorted(hipStream_t st, int kernel, const DeviceField &f,
                                const Scratch &s, int reset);
hipError_t launch_sort_pack_mixed(hipStream_t st, int kernel, const DeviceField &f, Scratch &s);
hipError_t launch_uj_mixed(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                           int reset);
hipError_t launch_uj_f64_targets(hipStream_t st, int kernel, const DeviceField &f,
                                 const Scratch &s, const double *tX, int64_t nt, double *U,
                                 double *J);
hipError_t launch_uj_mixed2_targets(hipStream_t st, int kernel, const DeviceField &f,
                                    const Scratch &s, const double4 *tgt, int64_t nt,
                                    const int32_t *tperm, int64_t ldo, double *U, double *J);
hipError_t launch_uj_mixed_targets(hipStream_t st, int kernel, const DeviceField &f,
                                   const Scratch &s, const double4 *tgt, int64_t nt,
                                   const int32_t *tperm, int64_t ldo, double *U, double *J,
                                   int reset);
// Estr into f.E; with DJ (9 x cap) also its core-size derivative into DE.
hipError_t launch_estr_f64(hipStream_t st, int kernel, int transposed, const DeviceField &f,
                           Scratch &s, const double *DJ = nullptr, double *DE = nullptr);
// Core-size derivatives D[U] (3 x cap), D[J] (9 x cap) of every particle from
// the records of the UJ that just ran (FP64 all-pairs / mixed near pairs).
hipError_t launch_dsig_f64(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                           double *DU, double *DJ);
hipError_t launch_dsig_mixed(hipStream_t st, int kernel, const DeviceField &f, const Scratch &s,
                             double *DU, double *DJ);
hipError_t launch_sfs_coeff(hipStream_t st, const DeviceField &f, double Cs, int clip);
// dynamic SFS: C from f.DJ(), f.DE(), Lagrangian average, bounds, clipping
hipError_t launch_dyn_coeff(hipStream_t st, const DeviceField &f, double alpha, double zeta0,
                            double rlxf, double minC, double maxC, int clip);
hipError_t launch_fill(hipStream_t st, double *p, int64_t n, double v);
hipError_t launch_update(hipStream_t st, const DeviceField &f, int transposed, int sfs,
                         double fcoef, double gcoef, double zeta0, const double Uinf[3],
                         double a, double b, double dt);
hipError_t launch_relax(hipStream_t st, const DeviceField &f, int kind, double rlxf);
hipError_t launch_diagnostics(hipStream_t st, const DeviceField &f, Scratch &s, double *out2);
size_t sort_temp_bytes(int64_t n);
hipError_t launch_bbox(hipStream_t st, const DeviceField &f, double *part, int *nparts);
// Mixed path split: Morton sort (perm), then packing of the sorted records.
hipError_t launch_sort_mixed(hipStream_t st, const DeviceField &f, Scratch &s);
hipError_t launch_pack_mixed(hipStream_t st, int kernel, const DeviceField &f, Scratch &s);
// Freestream per particle from a device table: uinf[3 fid[i] + k] (fid NULL:
// every particle reads uinf[0..2]).  Replaces the Uinf argument of launch_update.
hipError_t launch_update_dev(hipStream_t st, const DeviceField &f, int transposed, int sfs,
                             double fcoef, double gcoef, double zeta0, const double *uinf,
                             const int32_t *fid, double a, double b, double dt, int viscous,
                             double nu);

// Free-vortex force (force.hip): segments of the free field per bound vortex,
// then out[6 b + k] for nb bound vortices (bnd: 8 doubles each).
int ffv_segments(int64_t nb, int64_t nf);
hipError_t launch_ffv(hipStream_t st, int kernel, const double *bnd, int nb, const DeviceField &f,
                      int nseg, double *part, double *out);

// Viscous schemes (viscous.hip).  PSE rate into f.V; vorticity sums
// w = sum_j G_j zeta_s(x_i - x_j) with G of stride gld (s = sigma_j, or s0 > 0),
// static sources masked when mask_static; CG vector steps over 3 components
// (stride ld) with static entries held fixed.
hipError_t launch_pse(hipStream_t st, int kernel, const DeviceField &f, double nu, double4 *rec);
hipError_t launch_zeta_sum(hipStream_t st, int kernel, const DeviceField &f, const double *G,
                           int64_t gld, double s0, bool mask_static, double4 *rec, double *w,
                           int64_t wld);
hipError_t launch_cg_dot(hipStream_t st, const double *a, const double *b, const uint8_t *mask,
                         int64_t n, int64_t ld, double *part, double *out3);
hipError_t launch_cg_init(hipStream_t st, const double *w, const double *Ap, const uint8_t *mask,
                          int64_t n, int64_t ld, double *r, double *p);
hipError_t launch_cg_step(hipStream_t st, const double *coef, const double *p, const double *Ap,
                          const uint8_t *mask, int64_t n, int64_t ld, double *G, int64_t gld,
                          double *r);
hipError_t launch_cg_dir(hipStream_t st, const double *coef, const double *r, const uint8_t *mask,
                         int64_t n, int64_t ld, double *p);
hipError_t launch_max_sigma(hipStream_t st, const DeviceField &f, double *part, double *out);
hipError_t launch_set_sigma(hipStream_t st, const DeviceField &f, double s0);

// Batched fields (batch.hip).
size_t batch_sort_temp_bytes(int64_t n);
// per-field bbox, 64-bit keys (field << 30 | Morton), stable radix sort -> s.perm
hipError_t launch_batch_sort(hipStream_t st, const DeviceField &f, Scratch &s,
                             const BatchState &b);
// fills fid from the offsets
hipError_t launch_batch_fid(hipStream_t st, const BatchState &b, int64_t n);
// pairwise pass over the work list with the batch's schedule
// Bounds of the batch's per-field tiles (after the mixed pack).
hipError_t launch_batch_tile_bounds(hipStream_t st, const DeviceField &f, const Scratch &s,
                                    const BatchState &b);
hipError_t launch_batch_uj_mixed(hipStream_t st, int kernel, const DeviceField &f,
                                 const Scratch &s, BatchState &b, int reset, int sched,
                                 int block, int grid);
hipError_t launch_batch_uj_f64(hipStream_t st, int kernel, const DeviceField &f,
                               const Scratch &s, BatchState &b, int reset);
// resident workgroups per CU of the queue kernel for a block size
int batch_queue_occupancy(int kernel, int block);

double zeta0_of(int kernel);

}  // namespace vpmx
