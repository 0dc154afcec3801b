// vpmx.hip -- C ABI of libvpmx.so (declared in include/vpmx.h).
//
// Host side of the drop-in boundary: validates the caller's particle matrix
// and settings, gathers the fields into pinned SoA staging, keeps the field
// resident in HBM, sequences the kernels of one UJ evaluation or one time step
// on the context's stream and scatters results back.  No exception or abort
// crosses the ABI: every failure becomes a VPMX_E* code plus a message.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "device_math.h"
#include "fmm.h"
#include "internal.h"

using namespace vpmx;

// Scratch of vpmx_uj_device_noalloc: a fixed block in every device's global
// memory, held until the process ends, so that entry point allocates nothing.
// Host code never passes `scratch` itself (on the host it names a stand-in,
// not device memory): vpmx_create resolves each device's own address once,
// with that device current, into g_noalloc_ptr.
#define VPMX_NOALLOC_SCRATCH_BYTES (48u << 20)
static __device__ __align__(256) unsigned char scratch[VPMX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

thread_local std::string g_noctx_err;

int fail(vpmx_ctx *c, int code, const std::string &msg)
{
    if (c)
        c->err = msg;
    else
        g_noctx_err = msg;
    return code;
}

#define VPMX_HIP(ctx, expr)                                                               \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail((ctx), e_ == hipErrorOutOfMemory ? VPMX_ENOMEM : VPMX_EHIP,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));               \
    } while (0)

#define VPMX_TRY(expr)            \
    do {                          \
        int r_ = (expr);          \
        if (r_ != VPMX_OK) return r_; \
    } while (0)

int check_inject(vpmx_ctx *c)
{
    if (c->inject > 0) {
        c->inject--;
        return fail(c, VPMX_EINJECT, "injected fault");
    }
    return VPMX_OK;
}

// ---- timing -----------------------------------------------------------------
hipEvent_t take_event(Timer &t)
{
    if (!t.pool.empty()) {
        hipEvent_t e = t.pool.back();
        t.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// A named host-side range for tracers (rocprofv3 --marker-trace shows the
// library's phases on its timeline; a no-op when no tool is attached).  Not
// opened while a graph is being captured.
struct MarkScope {
    bool open = false;
    MarkScope(const vpmx_ctx *c, const char *name)
    {
        if (!c->capturing) open = roctxRangePush(name) >= 0;
    }
    ~MarkScope()
    {
        if (open) (void)roctxRangePop();
    }
};

struct TimedScope {
    vpmx_ctx *c;
    int kind;
    hipEvent_t a = nullptr;
    TimedScope(vpmx_ctx *c_, int kind_) : c(c_), kind(kind_)
    {
        if (c->timer.on && !c->capturing) {
            a = take_event(c->timer);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~TimedScope()
    {
        if (!a) return;
        hipEvent_t b = take_event(c->timer);
        if (!b) {
            c->timer.pool.push_back(a);
            return;
        }
        (void)hipEventRecord(b, c->stream);
        c->timer.pending.push_back({kind, a, b});
    }
};

void collect_timing(vpmx_ctx *c)
{
    for (auto &p : c->timer.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->timer.total_ms[p.kind] += ms;
            c->timer.count[p.kind] += 1;
        }
        c->timer.pool.push_back(p.a);
        c->timer.pool.push_back(p.b);
    }
    c->timer.pending.clear();
}

int sync(vpmx_ctx *c)
{
    VPMX_HIP(c, hipStreamSynchronize(c->stream));
    collect_timing(c);
    return VPMX_OK;
}

// ---- memory -------------------------------------------------------------------
template <class T>
int dalloc(vpmx_ctx *c, T **p, size_t count)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) return VPMX_OK;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess)
        return fail(c, VPMX_ENOMEM, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                        " B): " + hipGetErrorString(e));
    *p = (T *)q;
    return VPMX_OK;
}

void free_field(DeviceField &f)
{
    for (double **p : {&f.X, &f.G, &f.sigma, &f.U, &f.J, &f.E, &f.C, &f.M, &f.D, &f.vol, &f.V}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (f.is_static) (void)hipFree(f.is_static);
    f.is_static = nullptr;
    f.cap = 0;
    f.np = 0;
}

void free_scratch(Scratch &s)
{
    if (s.src64) (void)hipFree(s.src64);
    if (s.srcE) (void)hipFree(s.srcE);
    if (s.spos) (void)hipFree(s.spos);
    if (s.tbnd) (void)hipFree(s.tbnd);
    if (s.kpart) (void)hipFree(s.kpart);
    if (s.sgam) (void)hipFree(s.sgam);
    if (s.kin) (void)hipFree(s.kin);
    if (s.kout) (void)hipFree(s.kout);
    if (s.vin) (void)hipFree(s.vin);
    if (s.perm) (void)hipFree(s.perm);
    if (s.sort_tmp) (void)hipFree(s.sort_tmp);
    if (s.part) (void)hipFree(s.part);
    for (void *p : {(void *)s.tX, (void *)s.tU, (void *)s.tJ, (void *)s.tpos, (void *)s.tperm,
                    (void *)s.bnd, (void *)s.fpart, (void *)s.fout, (void *)s.vrec, (void *)s.cw,
                    (void *)s.cr, (void *)s.cp, (void *)s.cAp, (void *)s.cpart, (void *)s.cdev})
        if (p) (void)hipFree(p);
    s = Scratch();
}

constexpr int64_t PART_DOUBLES = 8 * 1024;   // bbox / reduction partials + result slots

int ensure_capacity(vpmx_ctx *c, int64_t np)
{
    DeviceField &f = c->f;
    Scratch &s = c->s;
    if (!s.part) VPMX_TRY(dalloc(c, &s.part, PART_DOUBLES));
    if (np <= f.cap) return VPMX_OK;
    int64_t cap = std::max<int64_t>({np, f.cap + f.cap / 2, 1024});
    free_field(f);
    s.vcap = 0;
    VPMX_TRY(dalloc(c, &f.X, 3 * cap));
    VPMX_TRY(dalloc(c, &f.G, 3 * cap));
    VPMX_TRY(dalloc(c, &f.sigma, cap));
    VPMX_TRY(dalloc(c, &f.U, 3 * cap));
    VPMX_TRY(dalloc(c, &f.J, 9 * cap));
    VPMX_TRY(dalloc(c, &f.E, 3 * cap));
    VPMX_TRY(dalloc(c, &f.C, 3 * cap));
    VPMX_TRY(dalloc(c, &f.D, 15 * cap));
    VPMX_TRY(dalloc(c, &f.M, 7 * cap));
    VPMX_TRY(dalloc(c, &f.vol, cap));
    VPMX_TRY(dalloc(c, &f.is_static, cap));
    f.cap = cap;
    VPMX_TRY(dalloc(c, &s.src64, 2 * cap));
    VPMX_TRY(dalloc(c, &s.srcE, 7 * cap));
    VPMX_TRY(dalloc(c, &s.spos, cap));
    VPMX_TRY(dalloc(c, &s.sgam, cap));
    VPMX_TRY(dalloc(c, &s.tbnd, 2 * ((cap + 255) / 256)));
    if (!s.kpart) {
        s.kcap = 15 * 256 * 1536;   // the most a split launch needs (see split_for)
        VPMX_TRY(dalloc(c, &s.kpart, s.kcap));
    }
    VPMX_TRY(dalloc(c, &s.kin, cap));
    VPMX_TRY(dalloc(c, &s.kout, cap));
    VPMX_TRY(dalloc(c, &s.vin, cap));
    VPMX_TRY(dalloc(c, &s.perm, cap));
    s.sort_tmp_bytes = sort_temp_bytes(cap);
    VPMX_TRY(dalloc(c, &s.sort_tmp, s.sort_tmp_bytes));
    s.cap = cap;
    c->gen++;
    return VPMX_OK;
}

// Buffers of the viscous schemes, sized with the field.
int ensure_viscous(vpmx_ctx *c)
{
    DeviceField &f = c->f;
    Scratch &s = c->s;
    if (f.V && s.vcap == f.cap) return VPMX_OK;
    const int64_t cap = f.cap;
    VPMX_TRY(dalloc(c, &f.V, 3 * cap));
    VPMX_TRY(dalloc(c, &s.vrec, 2 * cap));
    VPMX_TRY(dalloc(c, &s.cw, 3 * cap));
    VPMX_TRY(dalloc(c, &s.cr, 3 * cap));
    VPMX_TRY(dalloc(c, &s.cp, 3 * cap));
    VPMX_TRY(dalloc(c, &s.cAp, 3 * cap));
    if (!s.cpart) VPMX_TRY(dalloc(c, &s.cpart, 3 * 512));
    if (!s.cdev) VPMX_TRY(dalloc(c, &s.cdev, 16));
    s.vcap = cap;
    c->gen++;
    return VPMX_OK;
}

// Per-field tables (offsets, boxes, freestreams) for nf fields.
int ensure_batch_fields(vpmx_ctx *c, int64_t nf)
{
    BatchState &B = c->b;
    if (nf <= B.fcap) return VPMX_OK;
    const int64_t cap = std::max<int64_t>({nf, 2 * B.fcap, 16});
    VPMX_TRY(dalloc(c, &B.doff, cap + 1));
    VPMX_TRY(dalloc(c, &B.box, 6 * cap));
    VPMX_TRY(dalloc(c, &B.uinf, 3 * cap));
    B.fcap = cap;
    c->gen++;
    return VPMX_OK;
}

int ensure_batch(vpmx_ctx *c, int64_t nf, int64_t np)
{
    BatchState &B = c->b;
    VPMX_TRY(ensure_batch_fields(c, nf));
    if (!B.counter) VPMX_TRY(dalloc(c, &B.counter, 1));
    if (np <= B.pcap) return VPMX_OK;
    const int64_t cap = std::max<int64_t>({np, B.pcap + B.pcap / 2, 1024});
    VPMX_TRY(dalloc(c, &B.fid, cap));
    VPMX_TRY(dalloc(c, &B.kin, cap));
    VPMX_TRY(dalloc(c, &B.kout, cap));
    B.sort_tmp_bytes = batch_sort_temp_bytes(cap);
    VPMX_TRY(dalloc(c, &B.sort_tmp, B.sort_tmp_bytes));
    B.pcap = cap;
    c->gen++;
    return VPMX_OK;
}

void free_batch(BatchState &B)
{
    for (void *p : {(void *)B.doff, (void *)B.box, (void *)B.uinf, (void *)B.fid, (void *)B.kin,
                    (void *)B.kout, (void *)B.sort_tmp, (void *)B.items, (void *)B.counter,
                    (void *)B.tiles, (void *)B.tbnd})
        if (p) (void)hipFree(p);
    B = BatchState();
}

void drop_graph(GraphState &G)
{
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    G.exec = nullptr;
    G.key.clear();
}

// Work list of the batched pairwise pass for workgroups of `block` targets:
// one item per block of consecutive targets of one field, against all of the
// field's sources; fields in decreasing size (longest processing time first),
// blocks of a field in order.
int ensure_items(vpmx_ctx *c, int block)
{
    BatchState &B = c->b;
    if (B.items_block == block && B.items_off == B.off) return VPMX_OK;
    const int nf = B.nfields;
    std::vector<int> order(nf);
    for (int k = 0; k < nf; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return B.off[a + 1] - B.off[a] > B.off[b + 1] - B.off[b];
    });
    // per-field tiles of `block` slots, field by field (their bounds feed the
    // far-field tile test of the mixed pass)
    std::vector<int2> tiles;
    std::vector<int32_t> tbase(nf);
    for (int k = 0; k < nf; ++k) {
        tbase[k] = (int32_t)tiles.size();
        for (int64_t t = B.off[k]; t < B.off[k + 1]; t += block)
            tiles.push_back(make_int2((int)t, (int)std::min<int64_t>(t + block, B.off[k + 1])));
    }
    std::vector<BatchItem> items;
    for (int k : order) {
        const int64_t o = B.off[k], e = B.off[k + 1];
        for (int64_t t = o; t < e; t += block)
            items.push_back({(int32_t)t, (int32_t)std::min<int64_t>(t + block, e), (int32_t)o,
                             (int32_t)e, tbase[k], 0});
    }
    if ((int64_t)items.size() > B.icap) {
        const int64_t cap = std::max<int64_t>((int64_t)items.size(), 2 * B.icap);
        VPMX_TRY(dalloc(c, &B.items, cap));
        B.icap = cap;
        c->gen++;
    }
    if ((int64_t)tiles.size() > B.tcap) {
        const int64_t cap = std::max<int64_t>((int64_t)tiles.size(), 2 * B.tcap);
        VPMX_TRY(dalloc(c, &B.tiles, cap));
        VPMX_TRY(dalloc(c, &B.tbnd, 2 * cap));
        B.tcap = cap;
        c->gen++;
    }
    B.ntiles = (int64_t)tiles.size();
    if (!items.empty())
        VPMX_HIP(c, hipMemcpyAsync(B.items, items.data(), items.size() * sizeof(BatchItem),
                                   hipMemcpyHostToDevice, c->stream));
    if (!tiles.empty())
        VPMX_HIP(c, hipMemcpyAsync(B.tiles, tiles.data(), tiles.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, c->stream));
    VPMX_TRY(sync(c));
    B.nitems = (int64_t)items.size();
    B.items_block = block;
    B.items_off = B.off;
    return VPMX_OK;
}

int cu_count(vpmx_ctx *c)
{
    if (c->ncu <= 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || n <= 0)
            n = 1;
        c->ncu = n;
    }
    return c->ncu;
}

// Work list and queue grid for the resident batch, before anything is enqueued
// (the list upload synchronizes, so it never happens inside a capture).
int prepare_batch(vpmx_ctx *c, int kernel)
{
    BatchState &B = c->b;
    if (!B.nfields) return VPMX_OK;
    VPMX_TRY(ensure_items(c, B.block));
    if (B.grid > 0) {
        B.grid_eff = B.grid;
    } else {
        const int occ = batch_queue_occupancy(kernel, B.block);
        B.grid_eff = std::max(1, occ) * cu_count(c);
    }
    return VPMX_OK;
}

int ensure_stage(vpmx_ctx *c, size_t bytes)
{
    if (bytes <= c->h_stage_bytes) return VPMX_OK;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    c->h_stage_bytes = 0;
    size_t want = std::max(bytes, c->h_stage_bytes + c->h_stage_bytes / 2);
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(c, VPMX_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    c->h_stage = (double *)p;
    c->h_stage_bytes = want;
    return VPMX_OK;
}

// ---- validation -------------------------------------------------------------------
bool row_ok(int32_t off, int width, int32_t nrows, bool optional)
{
    if (off < 0) return optional && off == -1;
    return off + width <= nrows;
}

int check_layout(vpmx_ctx *c, const vpmx_layout *L)
{
    if (!L) return fail(c, VPMX_EINVAL, "layout is NULL");
    if (L->nrows <= 0) return fail(c, VPMX_EINVAL, "layout.nrows must be positive");
    struct { const char *name; int32_t off; int w; bool opt; } rows[] = {
        {"X", L->X, 3, false},         {"Gamma", L->Gamma, 3, false}, {"sigma", L->sigma, 1, false},
        {"vol", L->vol, 1, true},      {"U", L->U, 3, false},         {"J", L->J, 9, false},
        {"SFS", L->SFS, 3, true},      {"C", L->C, 3, true},          {"static", L->static_, 1, true}};
    for (auto &r : rows)
        if (!row_ok(r.off, r.w, L->nrows, r.opt))
            return fail(c, VPMX_EINVAL,
                        std::string("layout row ") + r.name + " = " + std::to_string(r.off) +
                            " out of range for nrows = " + std::to_string(L->nrows));
    return VPMX_OK;
}

int check_uj_params(vpmx_ctx *c, const vpmx_uj_params *P)
{
    if (!P) return fail(c, VPMX_EINVAL, "uj params is NULL");
    if (P->kernel < VPMX_KERNEL_SINGULAR || P->kernel > VPMX_KERNEL_WINCKELMANS)
        return fail(c, VPMX_EINVAL, "unknown kernel " + std::to_string(P->kernel));
    if (P->precision != VPMX_FP64 && P->precision != VPMX_MIXED)
        return fail(c, VPMX_EINVAL, "unknown precision " + std::to_string(P->precision));
    if (P->sfs != VPMX_SFS_NONE && P->sfs != VPMX_SFS_CONSTANT)
        return fail(c, VPMX_EINVAL, "unknown sfs model " + std::to_string(P->sfs));
    if (P->sfs && P->kernel == VPMX_KERNEL_SINGULAR)
        return fail(c, VPMX_EINVAL, "the SFS model needs a smooth kernel (zeta of singular is a delta)");
    if (P->fmm != 0 && P->fmm != 1)
        return fail(c, VPMX_EUNSUPPORTED, "fmm must be 0 (direct) or 1 (fast multipole)");
    if (P->fmm) {
        if (P->p < 0 || P->p > FMM_MAX_P)
            return fail(c, VPMX_EINVAL, "FMM expansion order p must be in [0, " +
                                            std::to_string(FMM_MAX_P) + "]");
        if (P->ncrit < 1) return fail(c, VPMX_EINVAL, "FMM ncrit must be positive");
        if (!(P->theta > 0.0) || !isfinite(P->theta))
            return fail(c, VPMX_EINVAL, "FMM theta must be positive and finite");
    }
    return VPMX_OK;
}

int check_step_params(vpmx_ctx *c, const vpmx_step_params *S, double dt)
{
    if (!S) return fail(c, VPMX_EINVAL, "step params is NULL");
    if (!isfinite(dt)) return fail(c, VPMX_EINVAL, "dt is not finite");
    if (S->relax_kind < 0 || S->relax_kind > 2)
        return fail(c, VPMX_EINVAL, "unknown relaxation kind " + std::to_string(S->relax_kind));
    if (S->sfs < VPMX_SFS_NONE || S->sfs > VPMX_SFS_DYNAMIC)
        return fail(c, VPMX_EINVAL, "unknown sfs model " + std::to_string(S->sfs));
    if (S->sfs == VPMX_SFS_DYNAMIC) {
        if (!(S->alpha > 0.0) || S->alpha == 1.0 || !isfinite(S->alpha))
            return fail(c, VPMX_EINVAL, "dynamic SFS needs a test filter factor alpha > 0, != 1");
        if (!(S->sfs_rlxf >= 0.0 && S->sfs_rlxf <= 1.0))
            return fail(c, VPMX_EINVAL, "dynamic SFS averaging weight must be in [0, 1]");
        if (!(S->minC <= S->maxC))
            return fail(c, VPMX_EINVAL, "dynamic SFS needs minC <= maxC");
    }
    if (!isfinite(S->f) || !isfinite(S->g) || 1.0 + 3.0 * S->f == 0.0)
        return fail(c, VPMX_EINVAL, "invalid formulation coefficients f, g");
    if (S->viscous < VPMX_VISCOUS_NONE || S->viscous > VPMX_VISCOUS_PSE)
        return fail(c, VPMX_EINVAL, "unknown viscous scheme " + std::to_string(S->viscous));
    if (S->viscous != VPMX_VISCOUS_NONE && !(S->nu >= 0.0 && isfinite(S->nu)))
        return fail(c, VPMX_EINVAL, "viscosity nu must be finite and >= 0");
    if (S->viscous == VPMX_VISCOUS_CORESPREADING) {
        if (!(S->sgm0 > 0.0) || !isfinite(S->sgm0))
            return fail(c, VPMX_EINVAL, "core spreading needs a reset core size sgm0 > 0");
        if (!(S->beta >= 1.0) || !isfinite(S->beta))
            return fail(c, VPMX_EINVAL, "core spreading needs a reset trigger beta >= 1");
        if (S->itmax < 0 || !(S->tol > 0.0))
            return fail(c, VPMX_EINVAL, "core spreading needs itmax >= 0 and tol > 0");
    }
    return VPMX_OK;
}

// ---- host <-> device ------------------------------------------------------------------
// Upload staging (doubles, stride = total np): X[3] G[3] sigma C[3] vol | U[3] J[9],
// followed by total np bytes of static flags.  Field k of a batch fills slots
// [off[k], off[k+1]).
std::string where(int nf, int k, int64_t p)
{
    std::string s = "particle " + std::to_string(p + 1);
    if (nf > 1) s = "field " + std::to_string(k + 1) + " " + s;
    return s;
}

int upload_fields(vpmx_ctx *c, int nf, const double *const *P, const int64_t *nps,
                  const vpmx_layout *L, bool with_uj)
{
    std::vector<int64_t> off(nf + 1, 0);
    for (int k = 0; k < nf; ++k) off[k + 1] = off[k] + nps[k];
    const int64_t np = off[nf];
    if (np > (int64_t)INT32_MAX) return fail(c, VPMX_EINVAL, "total np exceeds 2^31 - 1");
    DeviceField &f = c->f;
    BatchState &B = c->b;
    // until the upload completes the resident state is one empty field
    f.np = 0;
    B.nfields = 0;
    B.off.assign(2, 0);
    VPMX_TRY(ensure_capacity(c, np));
    const size_t nd = (size_t)(with_uj ? 23 : 11) * np;
    VPMX_TRY(ensure_stage(c, nd * sizeof(double) + np + 8));
    double *h = c->h_stage;
    uint8_t *hs = (uint8_t *)(h + nd);
    const int64_t ld = L->nrows;
    bool any_static = false;
    for (int fk = 0; fk < nf; ++fk) {
        for (int64_t p = 0; p < nps[fk]; ++p) {
            const double *col = P[fk] + p * ld;
            const int64_t q = off[fk] + p;
            for (int k = 0; k < 3; ++k) {
                const double x = col[L->X + k], g = col[L->Gamma + k];
                if (!isfinite(x) || !isfinite(g))
                    return fail(c, VPMX_EINVAL,
                                where(nf, fk, p) + " has a non-finite position or circulation");
                h[k * np + q] = x;
                h[(3 + k) * np + q] = g;
            }
            const double sg = col[L->sigma];
            if (!(sg > 0.0) || !isfinite(sg))
                return fail(c, VPMX_EINVAL, where(nf, fk, p) + " has sigma = " + std::to_string(sg) +
                                                " (must be positive and finite)");
            h[6 * np + q] = sg;
            for (int k = 0; k < 3; ++k) h[(7 + k) * np + q] = L->C >= 0 ? col[L->C + k] : 0.0;
            h[10 * np + q] = L->vol >= 0 ? col[L->vol] : sg * sg * sg;
            const uint8_t st = (L->static_ >= 0 && col[L->static_] != 0.0) ? 1 : 0;
            hs[q] = st;
            any_static |= st != 0;
            if (with_uj) {
                for (int k = 0; k < 3; ++k) h[(11 + k) * np + q] = col[L->U + k];
                for (int k = 0; k < 9; ++k) h[(14 + k) * np + q] = col[L->J + k];
            }
        }
    }
    f.np = np;
    f.any_static = any_static;
    B.nfields = nf > 1 ? nf : 0;
    B.off = off;
    if (np == 0) return VPMX_OK;
    hipStream_t st = c->stream;
    const size_t b = (size_t)np * sizeof(double);
    for (int k = 0; k < 3; ++k) {
        VPMX_HIP(c, hipMemcpyAsync(f.X + k * f.cap, h + k * np, b, hipMemcpyHostToDevice, st));
        VPMX_HIP(c, hipMemcpyAsync(f.G + k * f.cap, h + (3 + k) * np, b, hipMemcpyHostToDevice, st));
    }
    VPMX_HIP(c, hipMemcpyAsync(f.sigma, h + 6 * np, b, hipMemcpyHostToDevice, st));
    for (int k = 0; k < 3; ++k)
        VPMX_HIP(c, hipMemcpyAsync(f.C + k * f.cap, h + (7 + k) * np, b, hipMemcpyHostToDevice, st));
    VPMX_HIP(c, hipMemcpyAsync(f.vol, h + 10 * np, b, hipMemcpyHostToDevice, st));
    VPMX_HIP(c, hipMemcpyAsync(f.is_static, hs, (size_t)np, hipMemcpyHostToDevice, st));
    if (with_uj) {
        for (int k = 0; k < 3; ++k)
            VPMX_HIP(c, hipMemcpyAsync(f.U + k * f.cap, h + (11 + k) * np, b, hipMemcpyHostToDevice, st));
        for (int k = 0; k < 9; ++k)
            VPMX_HIP(c, hipMemcpyAsync(f.J + k * f.cap, h + (14 + k) * np, b, hipMemcpyHostToDevice, st));
    }
    if (B.nfields) {
        VPMX_TRY(ensure_batch(c, nf, np));
        VPMX_HIP(c, hipMemcpyAsync(B.doff, off.data(), off.size() * sizeof(int64_t),
                                   hipMemcpyHostToDevice, st));
        VPMX_HIP(c, launch_batch_fid(st, B, np));
    }
    return sync(c);   // staging is reused: the copies must finish
}

int upload(vpmx_ctx *c, const double *P, int64_t np, const vpmx_layout *L, bool with_uj)
{
    return upload_fields(c, 1, &P, &np, L, with_uj);
}

enum { DL_STATE = 1, DL_UJ = 2, DL_SFS = 4 };

int download_fields(vpmx_ctx *c, int nf, double *const *P, const vpmx_layout *L, int what)
{
    DeviceField &f = c->f;
    const int64_t np = f.np;
    if (np == 0) return VPMX_OK;
    const std::vector<int64_t> &off = c->b.off;
    VPMX_TRY(ensure_stage(c, (size_t)25 * np * sizeof(double)));
    double *h = c->h_stage;
    hipStream_t st = c->stream;
    const size_t b = (size_t)np * sizeof(double);
    if (what & DL_STATE) {
        for (int k = 0; k < 3; ++k) {
            VPMX_HIP(c, hipMemcpyAsync(h + k * np, f.X + k * f.cap, b, hipMemcpyDeviceToHost, st));
            VPMX_HIP(c, hipMemcpyAsync(h + (3 + k) * np, f.G + k * f.cap, b, hipMemcpyDeviceToHost, st));
        }
        VPMX_HIP(c, hipMemcpyAsync(h + 6 * np, f.sigma, b, hipMemcpyDeviceToHost, st));
    }
    if (what & DL_UJ) {
        for (int k = 0; k < 3; ++k)
            VPMX_HIP(c, hipMemcpyAsync(h + (7 + k) * np, f.U + k * f.cap, b, hipMemcpyDeviceToHost, st));
        for (int k = 0; k < 9; ++k)
            VPMX_HIP(c, hipMemcpyAsync(h + (10 + k) * np, f.J + k * f.cap, b, hipMemcpyDeviceToHost, st));
    }
    const bool sfs = (what & DL_SFS) != 0;
    if (sfs) {
        for (int k = 0; k < 3; ++k)
            VPMX_HIP(c, hipMemcpyAsync(h + (19 + k) * np, f.E + k * f.cap, b, hipMemcpyDeviceToHost, st));
        for (int k = 0; k < 3; ++k)
            VPMX_HIP(c, hipMemcpyAsync(h + (22 + k) * np, f.C + k * f.cap, b, hipMemcpyDeviceToHost, st));
    }
    VPMX_TRY(sync(c));
    const int64_t ld = L->nrows;
    for (int fk = 0; fk < nf; ++fk) {
        for (int64_t q = off[fk]; q < off[fk + 1]; ++q) {
            double *col = P[fk] + (q - off[fk]) * ld;
            if (what & DL_STATE) {
                for (int k = 0; k < 3; ++k) {
                    col[L->X + k] = h[k * np + q];
                    col[L->Gamma + k] = h[(3 + k) * np + q];
                }
                col[L->sigma] = h[6 * np + q];
            }
            if (what & DL_UJ) {
                for (int k = 0; k < 3; ++k) col[L->U + k] = h[(7 + k) * np + q];
                for (int k = 0; k < 9; ++k) col[L->J + k] = h[(10 + k) * np + q];
            }
            if (sfs) {
                if (L->SFS >= 0)
                    for (int k = 0; k < 3; ++k) col[L->SFS + k] = h[(19 + k) * np + q];
                if (L->C >= 0)
                    for (int k = 0; k < 3; ++k) col[L->C + k] = h[(22 + k) * np + q];
            }
        }
    }
    return VPMX_OK;
}

int download(vpmx_ctx *c, double *P, const vpmx_layout *L, int what)
{
    return download_fields(c, 1, &P, L, what);
}

// ---- fast multipole UJ ------------------------------------------------------------------
double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

double default_sigma_cut(int kernel)
{
    switch (kernel) {
    case VPMX_KERNEL_GAUSSIANERF: return 7.0;   // 1 - g(7) ~ 1e-10
    case VPMX_KERNEL_GAUSSIAN: return 3.0;      // e^{-27}
    case VPMX_KERNEL_WINCKELMANS: return 20.0;  // 1 - g ~ 15/(8 rho^4) ~ 1e-5
    default: return 0.0;
    }
}

template <class T>
int upload_vec(vpmx_ctx *c, T *dst, const std::vector<T> &v)
{
    if (v.empty()) return VPMX_OK;
    VPMX_HIP(c, hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return VPMX_OK;
}

template <class D, class S>
int upload_as(vpmx_ctx *c, D **dst, const std::vector<S> &v)
{
    const size_t bytes = v.size() * sizeof(S);
    VPMX_TRY(dalloc(c, (uint8_t **)dst, bytes > 0 ? bytes : 16));
    if (bytes) VPMX_HIP(c, hipMemcpy(*dst, v.data(), bytes, hipMemcpyHostToDevice));
    return VPMX_OK;
}

int fmm_tables_to_device(vpmx_ctx *c, FmmState &S)
{
    const FmmTables &T = S.tab;
    VPMX_TRY(upload_as(c, &S.mi, T.mi));
    VPMX_TRY(upload_as(c, &S.rec1, T.rec1));
    VPMX_TRY(upload_as(c, &S.rec2, T.rec2));
    VPMX_TRY(upload_as(c, &S.d1, T.d1));
    VPMX_TRY(upload_as(c, &S.d2a, T.d2a));
    VPMX_TRY(upload_as(c, &S.d2b, T.d2b));
    VPMX_TRY(upload_as(c, &S.m2m_off, T.m2m_off));
    VPMX_TRY(upload_as(c, &S.m2m_idx, T.m2m_idx));
    VPMX_TRY(upload_as(c, &S.m2m_coef, T.m2m_coef));
    VPMX_TRY(upload_as(c, &S.l2l_off, T.l2l_off));
    VPMX_TRY(upload_as(c, &S.l2l_idx, T.l2l_idx));
    VPMX_TRY(upload_as(c, &S.l2l_coef, T.l2l_coef));
    VPMX_TRY(upload_as(c, &S.m2l_toff, T.m2l_off));
    VPMX_TRY(upload_as(c, &S.m2l_tidx, T.m2l_idx));
    VPMX_TRY(upload_as(c, &S.m2l_tcoef, T.m2l_coef));
    VPMX_TRY(upload_as(c, &S.m2l_chunks, T.m2l_chunk));
    VPMX_TRY(upload_as(c, &S.m2l_cfirst, T.m2l_cfirst));
    VPMX_TRY(upload_as(c, &S.m2l_tile_idx, T.m2l_tile_idx));
    VPMX_TRY(upload_as(c, &S.m2l_bscale, T.m2l_bscale));
    VPMX_TRY(upload_as(c, &S.m2l_lscale, T.m2l_lscale));
    VPMX_TRY(upload_as(c, &S.m2l_erec, T.m2l_erec));
    S.dev_tab_p = T.P;
    return VPMX_OK;
}

void free_fmm(FmmState *S)
{
    if (!S) return;
    void *ptrs[] = {S->mi, S->m2m_off, S->l2l_off, S->m2l_toff, S->rec1, S->rec2, S->d1, S->d2a,
                    S->d2b, S->m2m_idx, S->l2l_idx, S->m2l_tidx, S->m2m_coef, S->l2l_coef,
                    S->m2l_tcoef, S->m2l_chunks, S->m2l_cfirst, S->m2l_tile_idx, S->m2l_bscale,
                    S->m2l_lscale, S->m2l_erec, S->chunk_buf, S->chunks, S->part, S->chunk_tmp,
                    S->cells, S->ijk, S->etab, S->leaves, S->parents, S->kids, S->parent,
                    S->m2l_off, S->p2p_off, S->m2l_src, S->p2p_src, S->M, S->L, S->kin, S->kout,
                    S->ssig, S->sort_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    const DttBufs &B = S->dtt;
    void *dptrs[] = {B.front, B.next, B.cnt_next, B.cnt_m2l, B.cnt_p2p, B.off_next, B.off_m2l,
                     B.off_p2p, B.m2l_keys, B.p2p_keys, B.keys_tmp, B.totals, B.tmp};
    for (void *p : dptrs)
        if (p) (void)hipFree(p);
    if (B.h_totals) (void)hipHostFree(B.h_totals);
    if (S->h_chunk_tot) (void)hipHostFree(S->h_chunk_tot);
    free_tree_bufs(S->tree);
    delete S;
}

int fmm_reserve_n(vpmx_ctx *c, FmmState &S, int64_t n)
{
    if (n <= S.cap_n) return VPMX_OK;
    const int64_t cap = std::max<int64_t>(n, S.cap_n + S.cap_n / 2);
    VPMX_TRY(dalloc(c, &S.kin, cap));
    VPMX_TRY(dalloc(c, &S.kout, cap));
    VPMX_TRY(dalloc(c, &S.ssig, cap));
    S.sort_bytes = fmm_sort_temp_bytes(cap);
    VPMX_TRY(dalloc(c, &S.sort_tmp, S.sort_bytes));
    S.cap_n = cap;
    return VPMX_OK;
}

// One FMM UJ evaluation; lists_only: the tree, the traversal and the
// near-field work items only (for the viscous sums), no evaluation.
int run_fmm(vpmx_ctx *c, const vpmx_uj_params *P, int reset, bool lists_only = false)
{
    MarkScope ms(c, lists_only ? "vpmx fmm lists" : "vpmx fmm uj");
    if (!c->fmm) c->fmm = new (std::nothrow) FmmState();
    if (!c->fmm) return fail(c, VPMX_ENOMEM, "out of host memory");
    FmmState &S = *c->fmm;
    DeviceField &f = c->f;
    const int64_t n = f.np;
    for (double &v : S.stats) v = 0.0;
    if (n == 0) return VPMX_OK;
    const double sigcut = P->sigma_cut < 0.0 ? default_sigma_cut(P->kernel) : P->sigma_cut;
    hipStream_t st = c->stream;

    // bounding cube (exact min/max, same arithmetic as the oracle)
    int nb = 0;
    VPMX_HIP(c, launch_bbox(st, f, c->s.part, &nb));
    std::vector<double> hp(6 * (size_t)nb);
    VPMX_HIP(c, hipMemcpyAsync(hp.data(), c->s.part, hp.size() * sizeof(double),
                               hipMemcpyDeviceToHost, st));
    VPMX_TRY(sync(c));
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = hp[k];
        hi[k] = hp[3 + k];
        for (int b = 1; b < nb; b++) {
            lo[k] = fmin(lo[k], hp[6 * b + k]);
            hi[k] = fmax(hi[k], hp[6 * b + 3 + k]);
        }
    }
    double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    if (!(ext > 0.0)) ext = 1.0;
    ext *= 1.0000001;
    const double scale = 2097152.0 / ext;

    // keys, sort, sorted records
    VPMX_TRY(fmm_reserve_n(c, S, n));
    VPMX_HIP(c, launch_fmm_keys(st, f, lo, scale, S.kin, c->s.vin));
    VPMX_HIP(c, launch_fmm_sort(st, S.sort_tmp, S.sort_bytes, S.kin, S.kout, c->s.vin, c->s.perm, n));
    VPMX_HIP(c, launch_fmm_gather(st, f, c->s.perm, c->s.src64, S.ssig));

    // device: octree of the sorted keys (one count read-back per level)
    FmmPlan &pl = S.plan;
    const double t0 = now_ms();
    for (int k = 0; k < 3; k++) pl.lo[k] = lo[k];
    pl.ext = ext;
    int64_t nc = 0;
    {
        MarkScope ms(c, "vpmx fmm tree");
        VPMX_HIP(c, fmm_build_tree_device(st, S, n, P->ncrit, pl, &nc));
    }
    // device: tables (once per p), expansions
    if (S.dev_tab_p != P->p) {
        fmm_build_tables(S.tab, P->p);
        VPMX_TRY(fmm_tables_to_device(c, S));
    }
    VPMX_HIP(c, fmm_m2l_table(st, S, P->p));
    const int K = S.tab.K;
    if ((nc + 1) * 3 * K > S.cap_coef) {   // + the zero row of the matrix-core M2L
        const int64_t cap = std::max<int64_t>((nc + 1) * 3 * K, S.cap_coef + S.cap_coef / 2);
        VPMX_TRY(dalloc(c, &S.M, cap));
        VPMX_TRY(dalloc(c, &S.L, cap));
        S.cap_coef = cap;
    }
    VPMX_TRY(sync(c));
    const double t1 = now_ms();
    const double t2 = t1;
    // device: dual tree traversal -> M2L and P2P lists (CSR by target)
    DttResult dr;
    {
        MarkScope ms(c, "vpmx fmm traversal");
        VPMX_HIP(c, fmm_traverse_device(st, S, nc, P->theta, sigcut, dr));
    }
    VPMX_HIP(c, fmm_p2p_chunks(st, S, pl.nleaves));
    const int64_t nm2l = dr.nm2l, np2p = dr.np2p, pairs = dr.p2p_pairs;
    const double t3 = now_ms();
    if (lists_only) {
        S.stats[0] = t1 - t0;
        S.stats[1] = t3 - t2;
        S.stats[2] = t2 - t1;
        S.stats[3] = (double)nc;
        S.stats[4] = (double)pl.nleaves;
        S.stats[5] = (double)nm2l;
        S.stats[6] = (double)np2p;
        S.stats[7] = (double)pairs;
        return VPMX_OK;
    }

    FmmDev d;
    d.P = P->p;
    d.K = K;
    d.ncell = (int)nc;
    d.nleaf = (int)pl.nleaves;
    d.cells = S.cells;
    d.leaves = S.leaves;
    d.parents = S.parents;
    d.kids = S.kids;
    d.parent = S.parent;
    d.m2l_off = S.m2l_off;
    d.p2p_off = S.p2p_off;
    d.m2l_src = S.m2l_src;
    d.p2p_src = S.p2p_src;
    d.rec = c->s.src64;
    d.perm = c->s.perm;
    d.M = S.M;
    d.L = S.L;
    d.mi = S.mi;
    d.rec1 = S.rec1;
    d.rec2 = S.rec2;
    d.d1 = S.d1;
    d.d2a = S.d2a;
    d.d2b = S.d2b;
    d.m2m_off = S.m2m_off;
    d.l2l_off = S.l2l_off;
    d.m2l_toff = S.m2l_toff;
    d.m2m_idx = S.m2m_idx;
    d.l2l_idx = S.l2l_idx;
    d.m2l_tidx = S.m2l_tidx;
    d.m2m_coef = S.m2m_coef;
    d.l2l_coef = S.l2l_coef;
    d.m2l_tcoef = S.m2l_tcoef;
    d.m2l_chunks = S.m2l_chunks;
    d.m2l_cfirst = S.m2l_cfirst;
    d.m2l_nchunk = S.tab.m2l_nchunk;
    d.m2l_rounds = S.tab.m2l_rounds;
    d.m2l_tile_idx = S.m2l_tile_idx;
    d.m2l_bscale = S.m2l_bscale;
    d.m2l_lscale = S.m2l_lscale;
    d.m2l_erec = S.m2l_erec;
    d.ijk = S.ijk;
    d.etab = S.etab;
    d.ext = pl.ext;
    d.chunks = S.chunks;
    d.chunk_off = S.chunk_off;
    d.part_base = S.part_base;
    d.part = S.part;
    d.nchunk = S.nchunk;
    d.nslot = S.nslot;
    const double ms_before = c->timer.total_ms[T_FMM];
    {
        TimedScope ts(c, T_FMM);
        MarkScope ms(c, "vpmx fmm evaluation");
        VPMX_HIP(c, launch_fmm_eval(st, d, pl, P->kernel, P->precision, f, reset));
    }
    VPMX_TRY(sync(c));
    S.stats[0] = t1 - t0;   // device tree build (wall, incl. its read-backs)
    S.stats[1] = t3 - t2;   // device traversal (incl. list sort)
    S.stats[2] = 0.0;       // no tree data crosses from the host
    S.stats[3] = (double)nc;
    S.stats[4] = (double)pl.nleaves;
    S.stats[5] = (double)nm2l;
    S.stats[6] = (double)np2p;
    S.stats[7] = (double)pairs;
    S.stats[8] = c->timer.total_ms[T_FMM] - ms_before;
    return VPMX_OK;
}

// ---- device sequences (enqueue only; callers sync) ---------------------------------
// Morton order of the mixed path (per field for a batch).
int enqueue_sort(vpmx_ctx *c)
{
    if (c->b.nfields)
        VPMX_HIP(c, launch_batch_sort(c->stream, c->f, c->s, c->b));
    else
        VPMX_HIP(c, launch_sort_mixed(c->stream, c->f, c->s));
    return VPMX_OK;
}

// One UJ evaluation.  `sort` = false reuses the Morton order of the last sort
// (a time step sorts once and keeps that order through its substeps).
int enqueue_uj(vpmx_ctx *c, const vpmx_uj_params *P, int reset, bool sort = true)
{
    const DeviceField &f = c->f;
    BatchState &B = c->b;
    if (P->fmm) {
        if (B.nfields) return fail(c, VPMX_EUNSUPPORTED, "the FMM does not run on batched fields");
        return run_fmm(c, P, reset);
    }
    MarkScope ms(c, "vpmx direct uj");
    if (B.nfields && (B.items_block != B.block || B.items_off != B.off))
        return fail(c, VPMX_EINVAL, "batch work list not prepared");
    if (P->precision == VPMX_FP64 && B.nfields) {
        VPMX_HIP(c, launch_pack_f64(c->stream, f, c->s));
        TimedScope ts(c, T_UJ);
        VPMX_HIP(c, launch_batch_uj_f64(c->stream, P->kernel, f, c->s, B, reset));
    } else if (P->precision == VPMX_FP64) {
        if (sort) VPMX_TRY(enqueue_sort(c));
        VPMX_HIP(c, launch_pack_f64_sorted(c->stream, f, c->s));
        TimedScope ts(c, T_UJ);
        VPMX_HIP(c, launch_uj_f64_sorted(c->stream, P->kernel, f, c->s, reset));
    } else {
        if (sort) VPMX_TRY(enqueue_sort(c));
        VPMX_HIP(c, launch_pack_mixed(c->stream, P->kernel, f, c->s));
        if (B.nfields) VPMX_HIP(c, launch_batch_tile_bounds(c->stream, f, c->s, B));
        TimedScope ts(c, T_UJ);
        if (B.nfields)
            VPMX_HIP(c, launch_batch_uj_mixed(c->stream, P->kernel, f, c->s, B, reset, B.sched,
                                              B.block, B.grid_eff));
        else
            VPMX_HIP(c, launch_uj_mixed(c->stream, P->kernel, f, c->s, reset));
    }
    return VPMX_OK;
}

// Estr with the scheme of the UJ that just ran (direct all-pairs, or the FMM's
// near-field lists); dyn: also its core-size derivative into f.DE() from f.DJ().
int enqueue_estr(vpmx_ctx *c, const vpmx_uj_params *U, bool dyn = false)
{
    TimedScope ts(c, T_ESTR);
    DeviceField &f = c->f;
    const double *DJ = dyn ? f.DJ() : nullptr;
    double *DE = dyn ? f.DE() : nullptr;
    if (U->fmm) {
        if (!c->fmm) return fail(c, VPMX_EINVAL, "Estr with the FMM needs an FMM UJ first");
        VPMX_HIP(c, launch_fmm_estr(c->stream, *c->fmm, U->kernel, U->transposed, U->precision, f,
                                    c->s, DJ, DE));
    } else {
        VPMX_HIP(c, launch_estr_f64(c->stream, U->kernel, U->transposed, f, c->s, DJ, DE));
    }
    return VPMX_OK;
}

// Dynamic procedure: core-size derivatives D[U], D[J] of the field, with the
// scheme and records of the stage's UJ (which must have just run).  The FMM's
// far field uses the singular kernel and does not depend on sigma, so with
// the FMM they are near-field sums over its P2P lists.
int enqueue_dsig(vpmx_ctx *c, const vpmx_uj_params *U)
{
    TimedScope ts(c, T_DSIG);
    DeviceField &f = c->f;
    if (U->fmm) {
        if (!c->fmm) return fail(c, VPMX_EINVAL, "the dynamic SFS with the FMM needs an FMM UJ first");
        VPMX_HIP(c, launch_fmm_dsig(c->stream, *c->fmm, U->kernel, U->precision, f, c->s, f.DU(),
                                    f.DJ()));
    } else if (U->precision == VPMX_FP64) {
        VPMX_HIP(c, launch_dsig_f64(c->stream, U->kernel, f, c->s, f.DU(), f.DJ()));
    } else {
        VPMX_HIP(c, launch_dsig_mixed(c->stream, U->kernel, f, c->s, f.DU(), f.DJ()));
    }
    return VPMX_OK;
}

// Low-storage RK3 coefficients (Williamson 1980), the scheme of
// docs/src/theory/rvpm.md:364.
constexpr double RK3_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double RK3_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

// The device work of one step after the step's sort: every substep's UJ (and
// Estr, SFS coefficient) and fused update, then the relaxation.  Enqueues
// kernels only, so it can be captured into a graph.  The freestream of every
// particle comes from the device table b.uinf (per field for a batch).
int enqueue_step(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S, double dt,
                 int relax, bool rk3)
{
    DeviceField &f = c->f;
    const int sfs = S->sfs != VPMX_SFS_NONE;
    const double zeta0 = zeta0_of(U->kernel);
    const int nst = rk3 ? 3 : 1;
    const int32_t *fid = c->b.nfields ? c->b.fid : nullptr;
    if (rk3) VPMX_HIP(c, launch_fill(c->stream, f.M, 7 * f.cap, 0.0));
    for (int s = 0; s < nst; ++s) {
        VPMX_TRY(enqueue_uj(c, U, 1, false));
        if (S->sfs == VPMX_SFS_DYNAMIC) {
            VPMX_TRY(enqueue_dsig(c, U));
            VPMX_TRY(enqueue_estr(c, U, true));
            VPMX_HIP(c, launch_dyn_coeff(c->stream, f, S->alpha, zeta0, S->sfs_rlxf, S->minC,
                                         S->maxC, S->clip_backscatter));
        } else if (sfs) {
            VPMX_TRY(enqueue_estr(c, U));
            VPMX_HIP(c, launch_sfs_coeff(c->stream, f, S->Cs, S->clip_backscatter));
        }
        if (S->viscous == VPMX_VISCOUS_PSE) {
            if (U->fmm)   // over the near lists of the stage's FMM evaluation
                VPMX_HIP(c, launch_fmm_pse(c->stream, *c->fmm, U->kernel, f, c->s, S->nu,
                                           c->s.vrec));
            else
                VPMX_HIP(c, launch_pse(c->stream, U->kernel, f, S->nu, c->s.vrec));
        }
        TimedScope ts(c, T_UPDATE);
        VPMX_HIP(c, launch_update_dev(c->stream, f, U->transposed, sfs, S->f, S->g, zeta0,
                                      c->b.uinf, fid, rk3 ? RK3_A[s] : 0.0,
                                      rk3 ? RK3_B[s] : 1.0, dt, S->viscous, S->nu));
    }
    if (relax && S->relax_kind != VPMX_RELAX_NONE) {
        if (rk3) VPMX_TRY(enqueue_uj(c, U, 1, false));
        VPMX_HIP(c, launch_relax(c->stream, f, S->relax_kind, S->rlxf));
    }
    return VPMX_OK;
}

// Core spreading's spatial adaptation on the resident field: when forced or a
// free core exceeds beta * sgm0, free cores are set to sgm0 and the free
// strengths re-fitted by conjugate gradients to the free particles' vorticity
// at the free particles (three independent systems, one per component).
// Host-driven: every iteration reads its three dot products back.
int core_spreading_reset(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S,
                         bool force, int *did, int *iters, double res[3])
{
    MarkScope ms(c, "vpmx core spreading");
    DeviceField &f = c->f;
    Scratch &s = c->s;
    hipStream_t st = c->stream;
    const int kernel = U->kernel;
    if (did) *did = 0;
    if (iters) *iters = 0;
    if (res) res[0] = res[1] = res[2] = 0.0;
    if (f.np == 0) return VPMX_OK;
    VPMX_TRY(ensure_viscous(c));
    if (!force) {
        double smax = 0.0;
        VPMX_HIP(c, launch_max_sigma(st, f, s.cpart, s.cdev));
        VPMX_HIP(c, hipMemcpyAsync(&smax, s.cdev, sizeof(double), hipMemcpyDeviceToHost, st));
        VPMX_TRY(sync(c));
        if (!(smax > S->beta * S->sgm0)) return VPMX_OK;
    }
    const int64_t n = f.np, ld = f.cap;
    const uint8_t *mask = f.any_static ? f.is_static : nullptr;
    // With the FMM every vorticity sum runs over the near lists of one tree
    // built here at the current positions and cores (zeta_fmm); directly
    // otherwise (zeta_direct).
    if (U->fmm) VPMX_TRY(run_fmm(c, U, 1, true));
    auto zeta = [&](const double *G, int64_t gld, double s0, double *w) -> hipError_t {
        if (U->fmm)
            return launch_fmm_zeta(st, *c->fmm, kernel, f, s, G, gld, s0, mask, s.vrec, w, ld);
        return launch_zeta_sum(st, kernel, f, G, gld, s0, true, s.vrec, w, ld);
    };
    // right-hand side: the free particles' vorticity with their current cores
    VPMX_HIP(c, zeta(f.G, f.cap, 0.0, s.cw));
    VPMX_HIP(c, launch_set_sigma(st, f, S->sgm0));
    VPMX_HIP(c, zeta(f.G, f.cap, S->sgm0, s.cAp));
    VPMX_HIP(c, launch_cg_init(st, s.cw, s.cAp, mask, n, ld, s.cr, s.cp));
    double h[6];
    VPMX_HIP(c, launch_cg_dot(st, s.cr, s.cr, mask, n, ld, s.cpart, s.cdev));
    VPMX_HIP(c, launch_cg_dot(st, s.cw, s.cw, mask, n, ld, s.cpart, s.cdev + 3));
    VPMX_HIP(c, hipMemcpyAsync(h, s.cdev, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
    VPMX_TRY(sync(c));
    double rr[3] = {h[0], h[1], h[2]}, ww[3] = {h[3], h[4], h[5]};
    bool done[3];
    for (int k = 0; k < 3; ++k) done[k] = !(sqrt(rr[k]) > S->tol * sqrt(ww[k]));
    int it = 0;
    while (it < S->itmax && !(done[0] && done[1] && done[2])) {
        VPMX_HIP(c, zeta(s.cp, ld, S->sgm0, s.cAp));
        double pAp[3];
        VPMX_HIP(c, launch_cg_dot(st, s.cp, s.cAp, mask, n, ld, s.cpart, s.cdev));
        VPMX_HIP(c, hipMemcpyAsync(pAp, s.cdev, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
        VPMX_TRY(sync(c));
        double coef[6];
        for (int k = 0; k < 3; ++k) coef[k] = done[k] || pAp[k] == 0.0 ? 0.0 : rr[k] / pAp[k];
        VPMX_HIP(c, hipMemcpyAsync(s.cdev + 8, coef, 3 * sizeof(double), hipMemcpyHostToDevice, st));
        VPMX_HIP(c, launch_cg_step(st, s.cdev + 8, s.cp, s.cAp, mask, n, ld, f.G, f.cap, s.cr));
        double rr1[3];
        VPMX_HIP(c, launch_cg_dot(st, s.cr, s.cr, mask, n, ld, s.cpart, s.cdev));
        VPMX_HIP(c, hipMemcpyAsync(rr1, s.cdev, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
        VPMX_TRY(sync(c));
        for (int k = 0; k < 3; ++k) {
            coef[k] = done[k] || rr[k] == 0.0 ? 0.0 : rr1[k] / rr[k];
            coef[3 + k] = done[k] ? 0.0 : 1.0;
        }
        VPMX_HIP(c, hipMemcpyAsync(s.cdev + 8, coef, 6 * sizeof(double), hipMemcpyHostToDevice, st));
        VPMX_HIP(c, launch_cg_dir(st, s.cdev + 8, s.cr, mask, n, ld, s.cp));
        for (int k = 0; k < 3; ++k) {
            rr[k] = rr1[k];
            done[k] = done[k] || !(sqrt(rr[k]) > S->tol * sqrt(ww[k]));
        }
        ++it;
    }
    VPMX_TRY(sync(c));
    if (did) *did = 1;
    if (iters) *iters = it;
    if (res)
        for (int k = 0; k < 3; ++k) res[k] = ww[k] > 0.0 ? sqrt(rr[k] / ww[k]) : 0.0;
    return VPMX_OK;
}

// Settings a captured step depends on, serialized field by field.
struct KeyWriter {
    std::vector<unsigned char> v;
    template <class T> void put(const T &x)
    {
        const unsigned char *p = (const unsigned char *)&x;
        v.insert(v.end(), p, p + sizeof(T));
    }
};

std::vector<unsigned char> step_key(vpmx_ctx *c, const vpmx_uj_params *U,
                                    const vpmx_step_params *S, double dt, int relax, bool rk3)
{
    KeyWriter k;
    k.put(U->kernel); k.put(U->precision); k.put(U->transposed); k.put(U->fmm);
    k.put(S->f); k.put(S->g); k.put(S->sfs); k.put(S->Cs); k.put(S->clip_backscatter);
    k.put(S->relax_kind); k.put(S->rlxf); k.put(S->alpha); k.put(S->sfs_rlxf);
    k.put(S->minC); k.put(S->maxC);
    k.put(S->viscous); k.put(S->nu);
    k.put(dt); k.put(relax); k.put(rk3);
    k.put(c->f.np); k.put(c->f.cap); k.put(c->f.any_static); k.put(c->gen);
    const BatchState &B = c->b;
    k.put(B.nfields); k.put(B.nitems); k.put(B.items_block); k.put(B.sched); k.put(B.grid_eff);
    return k.v;
}

// Runs one step on the resident state: per-step inputs (freestream table,
// Morton sort), then the step's device work, replayed from a graph when graphs
// are on and the step has no host-side phase (the FMM sizes each level of its
// tree and lists from a count read back to the host, so FMM steps run
// eagerly).  `uinf_fields` holds 3 values per field, or NULL for S->Uinf
// everywhere.
int run_step(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S,
             const double *uinf_fields, double dt, int relax, bool rk3)
{
    MarkScope ms(c, rk3 ? "vpmx rk3 step" : "vpmx euler step");
    BatchState &B = c->b;
    if (S->sfs != VPMX_SFS_NONE && U->kernel == VPMX_KERNEL_SINGULAR)
        return fail(c, VPMX_EINVAL, "the SFS model needs a smooth kernel");
    if (B.nfields && S->sfs != VPMX_SFS_NONE)
        return fail(c, VPMX_EUNSUPPORTED, "the SFS model does not run on batched fields");
    if (B.nfields && U->fmm)
        return fail(c, VPMX_EUNSUPPORTED, "the FMM does not run on batched fields");
    if (S->viscous != VPMX_VISCOUS_NONE) {
        if (B.nfields)
            return fail(c, VPMX_EUNSUPPORTED, "viscous schemes do not run on batched fields");
        if (U->kernel == VPMX_KERNEL_SINGULAR)
            return fail(c, VPMX_EINVAL, "viscous schemes need a smooth kernel");
        if (S->viscous == VPMX_VISCOUS_CORESPREADING && U->kernel != VPMX_KERNEL_GAUSSIANERF)
            return fail(c, VPMX_EINVAL,
                        "core spreading needs the gaussian-erf kernel (its cores solve the heat equation)");
    }
    if (c->f.np == 0) return VPMX_OK;
    if (S->viscous != VPMX_VISCOUS_NONE) VPMX_TRY(ensure_viscous(c));
    VPMX_TRY(prepare_batch(c, U->kernel));
    const int nf = std::max(1, B.nfields);
    VPMX_TRY(ensure_batch_fields(c, nf));
    if (c->h_small_bytes < (size_t)3 * nf * sizeof(double)) {
        if (c->h_small) (void)hipHostFree(c->h_small);
        c->h_small = nullptr;
        c->h_small_bytes = 0;
        void *p = nullptr;
        const size_t want = (size_t)3 * std::max(nf, 64) * sizeof(double);
        VPMX_HIP(c, hipHostMalloc(&p, want, hipHostMallocDefault));
        c->h_small = (double *)p;
        c->h_small_bytes = want;
    }
    for (int k = 0; k < nf; ++k)
        for (int j = 0; j < 3; ++j)
            c->h_small[3 * k + j] = uinf_fields ? uinf_fields[3 * k + j] : S->Uinf[j];
    for (int k = 0; k < 3 * nf; ++k)
        if (!isfinite(c->h_small[k])) return fail(c, VPMX_EINVAL, "freestream is not finite");
    VPMX_HIP(c, hipMemcpyAsync(B.uinf, c->h_small, (size_t)3 * nf * sizeof(double),
                               hipMemcpyHostToDevice, c->stream));
    if (!U->fmm && (U->precision == VPMX_MIXED || !B.nfields)) VPMX_TRY(enqueue_sort(c));
    if (S->viscous == VPMX_VISCOUS_CORESPREADING) {   // host-driven adaptation after the step
        VPMX_TRY(enqueue_step(c, U, S, dt, relax, rk3));
        return core_spreading_reset(c, U, S, false, nullptr, nullptr, nullptr);
    }
    if (!c->g.on || U->fmm) return enqueue_step(c, U, S, dt, relax, rk3);

    GraphState &G = c->g;
    std::vector<unsigned char> key = step_key(c, U, S, dt, relax, rk3);
    if (!G.exec || key != G.key) {
        drop_graph(G);
        VPMX_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        const int r = enqueue_step(c, U, S, dt, relax, rk3);
        c->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (r != VPMX_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess)
            return fail(c, VPMX_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            G.exec = nullptr;
            return fail(c, VPMX_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        G.key = key;
    }
    TimedScope ts(c, T_GRAPH);
    VPMX_HIP(c, hipGraphLaunch(G.exec, c->stream));
    return VPMX_OK;
}

int entry(vpmx_ctx *c)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    VPMX_TRY(check_inject(c));
    VPMX_HIP(c, hipSetDevice(c->device));
    return VPMX_OK;
}

int check_particles(vpmx_ctx *c, const double *P, int64_t np)
{
    if (np < 0) return fail(c, VPMX_EINVAL, "np is negative");
    if (np > 0 && !P) return fail(c, VPMX_EINVAL, "particles is NULL");
    if (np > (int64_t)INT32_MAX) return fail(c, VPMX_EINVAL, "np exceeds 2^31 - 1");
    return VPMX_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

const char *vpmx_version(void) { return "vpmx 0.1.0 (gfx950)"; }

int vpmx_create(vpmx_ctx **out, int32_t device)
{
    if (!out) return fail(nullptr, VPMX_EINVAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(nullptr, VPMX_ENODEV,
                    std::string("no HIP device available: ") + hipGetErrorString(e));
    if (device < 0 || device >= n)
        return fail(nullptr, VPMX_ENODEV, "device ordinal " + std::to_string(device) +
                                              " out of range (" + std::to_string(n) + " devices)");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, VPMX_ENODEV, hipGetErrorString(e));
    if (device < NOALLOC_MAX_DEVICES) {
        // once per device, with it current: its own address of `scratch`
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_noalloc_ptr[device]) {
            void *ptr = nullptr;
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr)
                return fail(nullptr, VPMX_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                                    hipGetErrorString(e));
            g_noalloc_ptr[device] = ptr;
        }
    }
    vpmx_ctx *c = new (std::nothrow) vpmx_ctx();
    if (!c) return fail(nullptr, VPMX_ENOMEM, "out of host memory");
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, VPMX_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return VPMX_OK;
}

void vpmx_destroy(vpmx_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->timer.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : c->timer.pool) (void)hipEventDestroy(e);
    drop_graph(c->g);
    free_field(c->f);
    free_scratch(c->s);
    free_batch(c->b);
    free_fmm(c->fmm);
    c->fmm = nullptr;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_small) (void)hipHostFree(c->h_small);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *vpmx_last_error(const vpmx_ctx *c) { return c ? c->err.c_str() : g_noctx_err.c_str(); }

int vpmx_uj(vpmx_ctx *c, double *particles, int64_t np, const vpmx_layout *L,
            const vpmx_uj_params *P)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_particles(c, particles, np));
    VPMX_TRY(check_layout(c, L));
    VPMX_TRY(check_uj_params(c, P));
    if (P->sfs && L->SFS < 0) return fail(c, VPMX_EINVAL, "sfs requested but layout.SFS = -1");
    VPMX_TRY(upload(c, particles, np, L, !P->reset));
    VPMX_TRY(enqueue_uj(c, P, P->reset));
    if (P->sfs) VPMX_TRY(enqueue_estr(c, P));
    VPMX_TRY(sync(c));
    return download(c, particles, L, DL_UJ | (P->sfs ? DL_SFS : 0));
}

int vpmx_uj_device_noalloc(vpmx_ctx *c, int64_t n, const double *X, const double *Gamma,
                           const double *sigma, double *U, double *J, const vpmx_uj_params *P,
                           void *stream)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, P));
    if (n < 0) return fail(c, VPMX_EINVAL, "n is negative");
    if (n > VPMX_NOALLOC_MAX_N)
        return fail(c, VPMX_EINVAL, "n = " + std::to_string(n) + " exceeds the no-allocation " +
                                        "limit of " + std::to_string(VPMX_NOALLOC_MAX_N) +
                                        " particles: use vpmx_uj");
    if (P->precision != VPMX_FP64 || P->fmm || P->sfs)
        return fail(c, VPMX_EUNSUPPORTED, "vpmx_uj_device_noalloc runs the FP64 direct UJ only "
                                          "(precision 0, fmm 0, sfs 0)");
    if (n == 0) return VPMX_OK;
    if (!X || !Gamma || !sigma || !U || !J)
        return fail(c, VPMX_EINVAL, "a device array is NULL");
    void *ptr = c->device < NOALLOC_MAX_DEVICES ? g_noalloc_ptr[c->device] : nullptr;
    if (!ptr)
        return fail(c, VPMX_EUNSUPPORTED, "no scratch block on device " +
                                              std::to_string(c->device));
    if (uj_f64_dev_scratch_bytes(n) > VPMX_NOALLOC_SCRATCH_BYTES)
        return fail(c, VPMX_EINVAL, "n = " + std::to_string(n) + " needs more scratch than the " +
                                        "fixed block holds");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    VPMX_HIP(c, launch_uj_f64_dev(st, P->kernel, n, X, Gamma, sigma, U, J, P->reset, ptr));
    return VPMX_OK;
}

static int step_host(vpmx_ctx *c, double *particles, int64_t np, const vpmx_layout *L,
                     const vpmx_uj_params *U, const vpmx_step_params *S, double dt, int32_t relax,
                     bool rk3)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_particles(c, particles, np));
    VPMX_TRY(check_layout(c, L));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(upload(c, particles, np, L, false));
    VPMX_TRY(run_step(c, U, S, nullptr, dt, relax, rk3));
    VPMX_TRY(sync(c));
    return download(c, particles, L, DL_STATE | DL_UJ | (S->sfs ? DL_SFS : 0));
}

int vpmx_rk3_step(vpmx_ctx *c, double *particles, int64_t np, const vpmx_layout *L,
                  const vpmx_uj_params *U, const vpmx_step_params *S, double dt, int32_t relax)
{
    return step_host(c, particles, np, L, U, S, dt, relax, true);
}

int vpmx_euler_step(vpmx_ctx *c, double *particles, int64_t np, const vpmx_layout *L,
                    const vpmx_uj_params *U, const vpmx_step_params *S, double dt, int32_t relax)
{
    return step_host(c, particles, np, L, U, S, dt, relax, false);
}

int vpmx_upload(vpmx_ctx *c, const double *particles, int64_t np, const vpmx_layout *L)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_particles(c, particles, np));
    VPMX_TRY(check_layout(c, L));
    return upload(c, particles, np, L, false);
}

int vpmx_download(vpmx_ctx *c, double *particles, int64_t np, const vpmx_layout *L)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_particles(c, particles, np));
    VPMX_TRY(check_layout(c, L));
    if (c->b.nfields)
        return fail(c, VPMX_EINVAL, "the resident state is a batch of " +
                                        std::to_string(c->b.nfields) +
                                        " fields: use vpmx_batch_download");
    if (np != c->f.np)
        return fail(c, VPMX_EINVAL, "np = " + std::to_string(np) + " but " +
                                        std::to_string(c->f.np) + " particles are resident");
    const int sfs = (L->SFS >= 0 || L->C >= 0) ? DL_SFS : 0;
    return download(c, particles, L, DL_STATE | DL_UJ | sfs);
}

int vpmx_uj_resident(vpmx_ctx *c, const vpmx_uj_params *P)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, P));
    if (c->b.nfields && P->sfs)
        return fail(c, VPMX_EUNSUPPORTED, "the SFS model does not run on batched fields");
    VPMX_TRY(prepare_batch(c, P->kernel));
    VPMX_TRY(enqueue_uj(c, P, P->reset));
    if (P->sfs) VPMX_TRY(enqueue_estr(c, P));
    return sync(c);
}

int vpmx_rk3_resident(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S, double dt,
                      int32_t relax)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(run_step(c, U, S, nullptr, dt, relax, true));
    return sync(c);
}

int vpmx_euler_resident(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S,
                        double dt, int32_t relax)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(run_step(c, U, S, nullptr, dt, relax, false));
    return sync(c);
}

int vpmx_diagnostics(vpmx_ctx *c, double *enstrophy, double *energy)
{
    VPMX_TRY(entry(c));
    if (!c->s.part) VPMX_TRY(ensure_capacity(c, 0));
    double *dout = c->s.part + (PART_DOUBLES - 8);
    double h[2] = {0.0, 0.0};
    if (c->f.np > 0) {
        VPMX_HIP(c, launch_diagnostics(c->stream, c->f, c->s, dout));
        VPMX_HIP(c, hipMemcpyAsync(h, dout, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        VPMX_TRY(sync(c));
    }
    if (enstrophy) *enstrophy = h[0];
    if (energy) *energy = h[1];
    return VPMX_OK;
}

int64_t vpmx_resident_np(const vpmx_ctx *c) { return c ? c->f.np : -1; }

int vpmx_set_timing(vpmx_ctx *c, int32_t on)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    c->timer.on = on != 0;
    for (int k = 0; k < T_NKIND; ++k) {
        c->timer.total_ms[k] = 0.0;
        c->timer.count[k] = 0;
    }
    return VPMX_OK;
}

int vpmx_get_timing(vpmx_ctx *c, int32_t which, double *total_ms, int64_t *count)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    if (which < 0 || which >= T_NKIND) return fail(c, VPMX_EINVAL, "unknown timing class");
    if (total_ms) *total_ms = c->timer.total_ms[which];
    if (count) *count = c->timer.count[which];
    return VPMX_OK;
}

int vpmx_uj_targets(vpmx_ctx *c, const double *targets, int64_t m, double *Uout, double *Jout,
                    const vpmx_uj_params *P)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, P));
    if (P->fmm) return fail(c, VPMX_EUNSUPPORTED, "target evaluation uses the direct scheme (fmm = 0)");
    if (P->sfs) return fail(c, VPMX_EUNSUPPORTED, "target evaluation has no SFS term");
    if (c->b.nfields)
        return fail(c, VPMX_EUNSUPPORTED, "target evaluation needs a single resident field");
    if (m < 0) return fail(c, VPMX_EINVAL, "number of targets is negative");
    if (m > (int64_t)INT32_MAX) return fail(c, VPMX_EINVAL, "number of targets exceeds 2^31 - 1");
    if (m > 0 && (!targets || !Uout || !Jout)) return fail(c, VPMX_EINVAL, "targets or outputs NULL");
    for (int64_t i = 0; i < 3 * m; i++)
        if (!isfinite(targets[i])) return fail(c, VPMX_EINVAL, "non-finite target position");
    if (m == 0) return VPMX_OK;
    const DeviceField &f = c->f;
    if (f.np == 0) {
        memset(Uout, 0, sizeof(double) * 3 * m);
        memset(Jout, 0, sizeof(double) * 9 * m);
        return VPMX_OK;
    }
    Scratch &s = c->s;
    if (m > s.tcap) {
        const int64_t cap = std::max<int64_t>(m, s.tcap + s.tcap / 2);
        VPMX_TRY(dalloc(c, &s.tX, 3 * cap));
        VPMX_TRY(dalloc(c, &s.tU, 3 * cap));
        VPMX_TRY(dalloc(c, &s.tJ, 9 * cap));
        VPMX_TRY(dalloc(c, &s.tpos, cap));
        VPMX_TRY(dalloc(c, &s.tperm, cap));
        s.tcap = cap;
    }
    VPMX_TRY(ensure_stage(c, (size_t)12 * m * sizeof(double)));
    hipStream_t st = c->stream;
    if (P->precision == VPMX_FP64) {
        double *h = c->h_stage;
        for (int64_t i = 0; i < m; i++)
            for (int k = 0; k < 3; k++) h[k * m + i] = targets[3 * i + k];
        VPMX_HIP(c, hipMemcpyAsync(s.tX, h, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
        VPMX_HIP(c, launch_pack_f64(st, f, s));
        TimedScope ts(c, T_UJ);
        VPMX_HIP(c, launch_uj_f64_targets(st, P->kernel, f, s, s.tX, m, s.tU, s.tJ));
    } else {
        // Morton order of the targets (30-bit keys in their bounding cube), so
        // each block's targets sit near the block origin.
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = 0; i < m; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = fmin(lo[k], targets[3 * i + k]);
                hi[k] = fmax(hi[k], targets[3 * i + k]);
            }
        const double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
        const double scale = ext > 0.0 ? 1023.999 / ext : 0.0;
        auto spread = [](uint32_t v) {
            v &= 0x3ffu;
            v = (v | (v << 16)) & 0x030000ffu;
            v = (v | (v << 8)) & 0x0300f00fu;
            v = (v | (v << 4)) & 0x030c30c3u;
            v = (v | (v << 2)) & 0x09249249u;
            return v;
        };
        std::vector<std::pair<uint32_t, int32_t>> ki(m);
        for (int64_t i = 0; i < m; i++) {
            uint32_t q[3];
            for (int k = 0; k < 3; k++)
                q[k] = (uint32_t)fmin(fmax((targets[3 * i + k] - lo[k]) * scale, 0.0), 1023.0);
            ki[i] = {spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2), (int32_t)i};
        }
        std::sort(ki.begin(), ki.end());
        double4 *hp = (double4 *)c->h_stage;
        int32_t *hperm = (int32_t *)(c->h_stage + 4 * m);
        for (int64_t j = 0; j < m; j++) {
            const int64_t i = ki[j].second;
            hp[j] = make_double4(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2], 0.0);
            hperm[j] = (int32_t)i;
        }
        VPMX_HIP(c, hipMemcpyAsync(s.tpos, hp, m * sizeof(double4), hipMemcpyHostToDevice, st));
        VPMX_HIP(c, hipMemcpyAsync(s.tperm, hperm, m * sizeof(int32_t), hipMemcpyHostToDevice, st));
        VPMX_HIP(c, launch_sort_pack_mixed(st, P->kernel, f, s));
        TimedScope ts(c, T_UJ);
        VPMX_HIP(c, launch_uj_mixed2_targets(st, P->kernel, f, s, s.tpos, m, s.tperm, m, s.tU,
                                             s.tJ));
    }
    double *h = c->h_stage;
    VPMX_HIP(c, hipMemcpyAsync(h, s.tU, 3 * m * sizeof(double), hipMemcpyDeviceToHost, st));
    VPMX_HIP(c, hipMemcpyAsync(h + 3 * m, s.tJ, 9 * m * sizeof(double), hipMemcpyDeviceToHost, st));
    VPMX_TRY(sync(c));
    for (int64_t i = 0; i < m; i++) {
        for (int k = 0; k < 3; k++) Uout[3 * i + k] = h[k * m + i];
        for (int k = 0; k < 9; k++) Jout[9 * i + k] = h[(3 + k) * m + i];
    }
    return VPMX_OK;
}

int vpmx_fmm_plan_counts(int64_t n, const double *X, const double *sigma, int32_t ncrit,
                         double theta, double sigma_cut, int64_t *out)
{
    if (n < 0 || (n > 0 && (!X || !sigma)) || !out || ncrit < 1)
        return fail(nullptr, VPMX_EINVAL, "bad arguments");
    FmmPlan pl;
    fmm_plan_host(pl, n, X, sigma, ncrit, theta, sigma_cut);
    int64_t nleaf = 0, pairs = 0;
    for (const auto &C : pl.cells) {
        if (C.nchild) continue;
        nleaf++;
    }
    for (size_t a = 0; a < pl.cells.size(); a++) {
        if (pl.cells[a].nchild) continue;
        int64_t s = 0;
        for (int64_t e = pl.p2p_off[a]; e < pl.p2p_off[a + 1]; e++)
            s += pl.cells[pl.p2p_src[e]].end - pl.cells[pl.p2p_src[e]].begin;
        pairs += s * (pl.cells[a].end - pl.cells[a].begin);
    }
    out[0] = (int64_t)pl.cells.size();
    out[1] = nleaf;
    out[2] = (int64_t)pl.m2l_src.size();
    out[3] = (int64_t)pl.p2p_src.size();
    out[4] = pairs;
    return VPMX_OK;
}

int vpmx_fmm_stats(vpmx_ctx *c, double *out, int32_t n)
{
    if (!c || !out) return fail(c, VPMX_EINVAL, "context or output is NULL");
    for (int k = 0; k < n && k < 9; k++) out[k] = c->fmm ? c->fmm->stats[k] : 0.0;
    return VPMX_OK;
}

void *vpmx_stream(vpmx_ctx *c) { return c ? (void *)c->stream : nullptr; }

int vpmx_inject_fault(vpmx_ctx *c, int32_t count)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    c->inject = count < 0 ? 0 : count;
    return VPMX_OK;
}

int vpmx_radial_f32(int32_t kernel, float r, float sigma, float *A, float *B)
{
    if (!A || !B || !(sigma > 0.f) || !(r > 0.f)) return VPMX_EINVAL;
    const float r2 = r * r, rinv = frsq(r2);
    float p0 = 0.f, p1 = 0.f;
    if (kernel == VPMX_KERNEL_GAUSSIAN || kernel == VPMX_KERNEL_GAUSSIANERF) {
        p0 = 1.0f / sigma;
        p1 = p0 * p0;
    } else if (kernel == VPMX_KERNEL_WINCKELMANS) {
        p0 = sigma * sigma;
    }
    switch (kernel) {
    case VPMX_KERNEL_SINGULAR: radial_near_f32<SINGULAR>(r2, rinv, p0, p1, *A, *B); break;
    case VPMX_KERNEL_GAUSSIAN: radial_near_f32<GAUSSIAN>(r2, rinv, p0, p1, *A, *B); break;
    case VPMX_KERNEL_GAUSSIANERF: radial_near_f32<GAUSSIANERF>(r2, rinv, p0, p1, *A, *B); break;
    case VPMX_KERNEL_WINCKELMANS: radial_near_f32<WINCKELMANS>(r2, rinv, p0, p1, *A, *B); break;
    default: return VPMX_EINVAL;
    }
    return VPMX_OK;
}

// ---- batched fields ------------------------------------------------------------
static int check_batch(vpmx_ctx *c, int32_t nf, const void *P, const int64_t *np)
{
    if (nf < 1) return fail(c, VPMX_EINVAL, "nfields must be at least 1");
    if (!P || !np) return fail(c, VPMX_EINVAL, "particles or np is NULL");
    int64_t tot = 0;
    for (int32_t k = 0; k < nf; ++k) {
        if (np[k] < 0) return fail(c, VPMX_EINVAL, "np of field " + std::to_string(k + 1) + " is negative");
        if (np[k] > 0 && !((const double *const *)P)[k])
            return fail(c, VPMX_EINVAL, "particles of field " + std::to_string(k + 1) + " is NULL");
        tot += np[k];
        if (tot > (int64_t)INT32_MAX) return fail(c, VPMX_EINVAL, "total np exceeds 2^31 - 1");
    }
    return VPMX_OK;
}

static int check_batch_uj(vpmx_ctx *c, int32_t nf, const vpmx_uj_params *P)
{
    if (nf > 1 && P->fmm) return fail(c, VPMX_EUNSUPPORTED, "the FMM does not run on batched fields");
    if (nf > 1 && P->sfs)
        return fail(c, VPMX_EUNSUPPORTED, "the SFS model does not run on batched fields");
    return VPMX_OK;
}

int vpmx_batch_upload(vpmx_ctx *c, int32_t nfields, const double *const *particles,
                      const int64_t *np, const vpmx_layout *L)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_batch(c, nfields, particles, np));
    VPMX_TRY(check_layout(c, L));
    return upload_fields(c, nfields, particles, np, L, false);
}

int vpmx_batch_download(vpmx_ctx *c, int32_t nfields, double *const *particles,
                        const int64_t *np, const vpmx_layout *L)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_batch(c, nfields, particles, np));
    VPMX_TRY(check_layout(c, L));
    const int32_t nres = vpmx_batch_nfields(c);
    if (nfields != nres)
        return fail(c, VPMX_EINVAL, "nfields = " + std::to_string(nfields) + " but " +
                                        std::to_string(nres) + " fields are resident");
    for (int32_t k = 0; k < nfields; ++k)
        if (np[k] != c->b.off[k + 1] - c->b.off[k])
            return fail(c, VPMX_EINVAL, "np of field " + std::to_string(k + 1) +
                                            " does not match the resident field");
    const int sfs = (L->SFS >= 0 || L->C >= 0) ? DL_SFS : 0;
    return download_fields(c, nfields, particles, L, DL_STATE | DL_UJ | sfs);
}

int vpmx_batch_uj(vpmx_ctx *c, int32_t nfields, double *const *particles, const int64_t *np,
                  const vpmx_layout *L, const vpmx_uj_params *P)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_batch(c, nfields, particles, np));
    VPMX_TRY(check_layout(c, L));
    VPMX_TRY(check_uj_params(c, P));
    VPMX_TRY(check_batch_uj(c, nfields, P));
    if (P->sfs && L->SFS < 0) return fail(c, VPMX_EINVAL, "sfs requested but layout.SFS = -1");
    VPMX_TRY(upload_fields(c, nfields, (const double *const *)particles, np, L, !P->reset));
    VPMX_TRY(prepare_batch(c, P->kernel));
    VPMX_TRY(enqueue_uj(c, P, P->reset));
    if (P->sfs) VPMX_TRY(enqueue_estr(c, P));
    VPMX_TRY(sync(c));
    return download_fields(c, nfields, particles, L, DL_UJ | (P->sfs ? DL_SFS : 0));
}

int vpmx_batch_rk3_step(vpmx_ctx *c, int32_t nfields, double *const *particles,
                        const int64_t *np, const vpmx_layout *L, const vpmx_uj_params *U,
                        const vpmx_step_params *S, const double *Uinf, double dt, int32_t relax)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_batch(c, nfields, particles, np));
    VPMX_TRY(check_layout(c, L));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(check_batch_uj(c, nfields, U));
    VPMX_TRY(upload_fields(c, nfields, (const double *const *)particles, np, L, false));
    VPMX_TRY(run_step(c, U, S, Uinf, dt, relax, true));
    VPMX_TRY(sync(c));
    return download_fields(c, nfields, particles, L, DL_STATE | DL_UJ | (S->sfs ? DL_SFS : 0));
}

int vpmx_batch_rk3_resident(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S,
                            const double *Uinf, double dt, int32_t relax)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(run_step(c, U, S, Uinf, dt, relax, true));
    return sync(c);
}

int32_t vpmx_batch_nfields(const vpmx_ctx *c)
{
    if (!c) return -1;
    if (c->b.nfields) return c->b.nfields;
    return c->b.off.size() > 1 ? 1 : 0;
}

int vpmx_batch_set_schedule(vpmx_ctx *c, int32_t schedule, int32_t block, int32_t grid)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    if (schedule != VPMX_SCHED_FLAT && schedule != VPMX_SCHED_QUEUE)
        return fail(c, VPMX_EINVAL, "unknown schedule " + std::to_string(schedule));
    if (block != 64 && block != 128 && block != 256)
        return fail(c, VPMX_EINVAL, "block must be 64, 128 or 256");
    if (grid < 0) return fail(c, VPMX_EINVAL, "grid must be >= 0");
    c->b.sched = schedule;
    c->b.block = block;
    c->b.grid = grid;
    c->b.tuned = false;
    return VPMX_OK;
}

int vpmx_batch_get_schedule(vpmx_ctx *c, int32_t *schedule, int32_t *block, int32_t *grid)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    if (schedule) *schedule = c->b.sched;
    if (block) *block = c->b.block;
    if (grid) *grid = c->b.grid > 0 ? c->b.grid : c->b.grid_eff;
    return VPMX_OK;
}

int vpmx_batch_autotune(vpmx_ctx *c, const vpmx_uj_params *P, int32_t reps, double *ms,
                        int32_t nms, int32_t *ncand)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, P));
    if (P->fmm || P->sfs)
        return fail(c, VPMX_EUNSUPPORTED, "the autotuner covers the direct UJ without SFS");
    BatchState &B = c->b;
    if (!B.nfields) return fail(c, VPMX_EINVAL, "no batch of several fields is resident");
    if (nms > 0 && !ms) return fail(c, VPMX_EINVAL, "ms is NULL");
    if (reps < 1) reps = 1;
    const bool mixed = P->precision == VPMX_MIXED;
    struct Cand { int sched, block; };
    std::vector<Cand> cands;
    for (int blk : {64, 128, 256}) {
        cands.push_back({SCHED_FLAT, blk});
        if (mixed) cands.push_back({SCHED_QUEUE, blk});
    }
    if (ncand) *ncand = (int32_t)cands.size();
    if (mixed) {
        VPMX_TRY(enqueue_sort(c));
        VPMX_HIP(c, launch_pack_mixed(c->stream, P->kernel, c->f, c->s));
    } else {
        VPMX_HIP(c, launch_pack_f64(c->stream, c->f, c->s));
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VPMX_HIP(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(c, VPMX_EHIP, "hipEventCreate failed");
    }
    auto launch = [&]() -> hipError_t {
        if (!mixed) return launch_batch_uj_f64(c->stream, P->kernel, c->f, c->s, B, 1);
        const hipError_t e = launch_batch_tile_bounds(c->stream, c->f, c->s, B);
        return e != hipSuccess ? e
                               : launch_batch_uj_mixed(c->stream, P->kernel, c->f, c->s, B, 1,
                                                       B.sched, B.block, B.grid_eff);
    };
    int rc = VPMX_OK, best = -1;
    double best_ms = INFINITY;
    for (size_t i = 0; i < cands.size() && rc == VPMX_OK; ++i) {
        B.sched = cands[i].sched;
        B.block = cands[i].block;
        B.grid = 0;
        rc = prepare_batch(c, P->kernel);
        if (rc != VPMX_OK) break;
        if (i == 0 && launch() != hipSuccess) {   // warm-up: first-launch costs
            rc = fail(c, VPMX_EHIP, "batched UJ launch failed");
            break;
        }
        double tmin = INFINITY;
        for (int r = 0; r < reps; ++r) {
            hipError_t e = hipEventRecord(e0, c->stream);
            if (e == hipSuccess) e = launch();
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e != hipSuccess) {
                rc = fail(c, VPMX_EHIP, std::string("autotune: ") + hipGetErrorString(e));
                break;
            }
            tmin = std::min(tmin, (double)t);
        }
        if (i < (size_t)std::max(nms, 0)) ms[i] = tmin;
        if (tmin < best_ms) {
            best_ms = tmin;
            best = (int)i;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    VPMX_TRY(rc);
    B.sched = cands[best].sched;
    B.block = cands[best].block;
    B.grid = 0;
    VPMX_TRY(prepare_batch(c, P->kernel));
    B.tuned = true;
    return sync(c);
}

int vpmx_set_graph(vpmx_ctx *c, int32_t on)
{
    if (!c) return fail(nullptr, VPMX_EINVAL, "context is NULL");
    c->g.on = on != 0;
    if (!c->g.on) {
        (void)hipSetDevice(c->device);
        drop_graph(c->g);
    }
    return VPMX_OK;
}

// ---- free-vortex force ------------------------------------------------------------
int vpmx_ffv(vpmx_ctx *c, const double *bound, int64_t nb, const vpmx_layout *L, int32_t kernel,
             double *out)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_particles(c, bound, nb));
    VPMX_TRY(check_layout(c, L));
    if (kernel < VPMX_KERNEL_SINGULAR || kernel > VPMX_KERNEL_WINCKELMANS)
        return fail(c, VPMX_EINVAL, "unknown kernel " + std::to_string(kernel));
    if (nb > 0 && !out) return fail(c, VPMX_EINVAL, "out is NULL");
    if (nb == 0) return VPMX_OK;
    if (c->b.nfields)
        return fail(c, VPMX_EUNSUPPORTED, "the free-vortex force needs a single resident field");
    std::vector<double> h(8 * (size_t)nb, 0.0);
    for (int64_t b = 0; b < nb; ++b) {
        const double *col = bound + b * L->nrows;
        for (int k = 0; k < 3; ++k) {
            h[8 * b + k] = col[L->X + k];
            h[8 * b + 3 + k] = col[L->Gamma + k];
        }
        const double sg = col[L->sigma];
        if (!(sg > 0.0) || !isfinite(sg))
            return fail(c, VPMX_EINVAL, "bound particle " + std::to_string(b + 1) +
                                            " has sigma = " + std::to_string(sg));
        for (int k = 0; k < 6; ++k)
            if (!isfinite(h[8 * b + k]))
                return fail(c, VPMX_EINVAL, "bound particle " + std::to_string(b + 1) +
                                                " has a non-finite position or circulation");
        h[8 * b + 6] = 1.0 / sg;
    }
    Scratch &s = c->s;
    if (!s.part) VPMX_TRY(ensure_capacity(c, 0));
    const int nseg = ffv_segments(nb, c->f.np);
    if (nb > s.bcap) {
        const int64_t cap = std::max<int64_t>(nb, 2 * s.bcap);
        VPMX_TRY(dalloc(c, &s.bnd, 8 * cap));
        VPMX_TRY(dalloc(c, &s.fout, 6 * cap));
        s.bcap = cap;
    }
    if (nb * nseg > s.fpcap) {
        const int64_t cap = std::max<int64_t>(nb * nseg, 2 * s.fpcap);
        VPMX_TRY(dalloc(c, &s.fpart, 6 * cap));
        s.fpcap = cap;
    }
    VPMX_HIP(c, hipMemcpyAsync(s.bnd, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
    VPMX_HIP(c, launch_ffv(c->stream, kernel, s.bnd, (int)nb, c->f, nseg, s.fpart, s.fout));
    VPMX_HIP(c, hipMemcpyAsync(out, s.fout, 6 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    return sync(c);
}

// ---- viscous: core spreading adaptation ------------------------------------------
int vpmx_core_spreading_reset(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S,
                              int32_t force, int32_t *did, int32_t *iters, double *res)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, 0.0));
    if (S->viscous != VPMX_VISCOUS_CORESPREADING)
        return fail(c, VPMX_EINVAL, "step params do not select core spreading");
    if (U->kernel != VPMX_KERNEL_GAUSSIANERF)
        return fail(c, VPMX_EINVAL, "core spreading needs the gaussian-erf kernel");
    if (c->b.nfields)
        return fail(c, VPMX_EUNSUPPORTED, "viscous schemes do not run on batched fields");
    int d = 0, it = 0;
    double r3[3];
    VPMX_TRY(core_spreading_reset(c, U, S, force != 0, &d, &it, r3));
    if (did) *did = d;
    if (iters) *iters = it;
    if (res)
        for (int k = 0; k < 3; ++k) res[k] = r3[k];
    return VPMX_OK;
}

}  // extern "C"
