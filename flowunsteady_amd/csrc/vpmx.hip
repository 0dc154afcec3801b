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
#include <new>
#include <string>

#include "device_math.h"
#include "internal.h"

using namespace vpmx;

namespace {

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

struct TimedScope {
    vpmx_ctx *c;
    int kind;
    hipEvent_t a = nullptr;
    TimedScope(vpmx_ctx *c_, int kind_) : c(c_), kind(kind_)
    {
        if (c->timer.on) {
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
    for (double **p : {&f.X, &f.G, &f.sigma, &f.U, &f.J, &f.E, &f.C, &f.M}) {
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
    if (s.sgam) (void)hipFree(s.sgam);
    if (s.kin) (void)hipFree(s.kin);
    if (s.kout) (void)hipFree(s.kout);
    if (s.vin) (void)hipFree(s.vin);
    if (s.perm) (void)hipFree(s.perm);
    if (s.sort_tmp) (void)hipFree(s.sort_tmp);
    if (s.part) (void)hipFree(s.part);
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
    VPMX_TRY(dalloc(c, &f.X, 3 * cap));
    VPMX_TRY(dalloc(c, &f.G, 3 * cap));
    VPMX_TRY(dalloc(c, &f.sigma, cap));
    VPMX_TRY(dalloc(c, &f.U, 3 * cap));
    VPMX_TRY(dalloc(c, &f.J, 9 * cap));
    VPMX_TRY(dalloc(c, &f.E, 3 * cap));
    VPMX_TRY(dalloc(c, &f.C, cap));
    VPMX_TRY(dalloc(c, &f.M, 7 * cap));
    VPMX_TRY(dalloc(c, &f.is_static, cap));
    f.cap = cap;
    VPMX_TRY(dalloc(c, &s.src64, 2 * cap));
    VPMX_TRY(dalloc(c, &s.srcE, 5 * cap));
    VPMX_TRY(dalloc(c, &s.spos, cap));
    VPMX_TRY(dalloc(c, &s.sgam, cap));
    VPMX_TRY(dalloc(c, &s.kin, cap));
    VPMX_TRY(dalloc(c, &s.kout, cap));
    VPMX_TRY(dalloc(c, &s.vin, cap));
    VPMX_TRY(dalloc(c, &s.perm, cap));
    s.sort_tmp_bytes = sort_temp_bytes(cap);
    VPMX_TRY(dalloc(c, &s.sort_tmp, s.sort_tmp_bytes));
    s.cap = cap;
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
        {"SFS", L->SFS, 3, true},      {"C", L->C, 1, true},          {"static", L->static_, 1, true}};
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
    if (P->fmm != 0)
        return fail(c, VPMX_EUNSUPPORTED, "fmm != 0: only the direct scheme is built");
    return VPMX_OK;
}

int check_step_params(vpmx_ctx *c, const vpmx_step_params *S, double dt)
{
    if (!S) return fail(c, VPMX_EINVAL, "step params is NULL");
    if (!isfinite(dt)) return fail(c, VPMX_EINVAL, "dt is not finite");
    if (S->relax_kind < 0 || S->relax_kind > 2)
        return fail(c, VPMX_EINVAL, "unknown relaxation kind " + std::to_string(S->relax_kind));
    if (S->sfs != VPMX_SFS_NONE && S->sfs != VPMX_SFS_CONSTANT)
        return fail(c, VPMX_EINVAL, "unknown sfs model " + std::to_string(S->sfs));
    if (!isfinite(S->f) || !isfinite(S->g) || 1.0 + 3.0 * S->f == 0.0)
        return fail(c, VPMX_EINVAL, "invalid formulation coefficients f, g");
    return VPMX_OK;
}

// ---- host <-> device ------------------------------------------------------------------
// Staging layout (doubles): X[3 np] G[3 np] sigma[np] | U[3 np] J[9 np] | E[3 np] C[np]
// followed by np bytes of static flags.
int upload(vpmx_ctx *c, const double *P, int64_t np, const vpmx_layout *L, bool with_uj)
{
    VPMX_TRY(ensure_capacity(c, np));
    DeviceField &f = c->f;
    f.np = np;
    if (np == 0) return VPMX_OK;
    const size_t nd = (size_t)(with_uj ? 19 : 7) * np;
    VPMX_TRY(ensure_stage(c, nd * sizeof(double) + np));
    double *h = c->h_stage;
    uint8_t *hs = (uint8_t *)(h + nd);
    const int64_t ld = L->nrows;
    bool any_static = false;
    for (int64_t p = 0; p < np; ++p) {
        const double *col = P + p * ld;
        for (int k = 0; k < 3; ++k) {
            const double x = col[L->X + k], g = col[L->Gamma + k];
            if (!isfinite(x) || !isfinite(g))
                return fail(c, VPMX_EINVAL, "particle " + std::to_string(p + 1) +
                                                " has a non-finite position or circulation");
            h[k * np + p] = x;
            h[(3 + k) * np + p] = g;
        }
        const double s = col[L->sigma];
        if (!(s > 0.0) || !isfinite(s))
            return fail(c, VPMX_EINVAL, "particle " + std::to_string(p + 1) + " has sigma = " +
                                            std::to_string(s) + " (must be positive and finite)");
        h[6 * np + p] = s;
        const uint8_t st = (L->static_ >= 0 && col[L->static_] != 0.0) ? 1 : 0;
        hs[p] = st;
        any_static |= st != 0;
        if (with_uj) {
            for (int k = 0; k < 3; ++k) h[(7 + k) * np + p] = col[L->U + k];
            for (int k = 0; k < 9; ++k) h[(10 + k) * np + p] = col[L->J + k];
        }
    }
    f.any_static = any_static;
    hipStream_t st = c->stream;
    const size_t b = (size_t)np * sizeof(double);
    for (int k = 0; k < 3; ++k) {
        VPMX_HIP(c, hipMemcpyAsync(f.X + k * f.cap, h + k * np, b, hipMemcpyHostToDevice, st));
        VPMX_HIP(c, hipMemcpyAsync(f.G + k * f.cap, h + (3 + k) * np, b, hipMemcpyHostToDevice, st));
    }
    VPMX_HIP(c, hipMemcpyAsync(f.sigma, h + 6 * np, b, hipMemcpyHostToDevice, st));
    VPMX_HIP(c, hipMemcpyAsync(f.is_static, hs, (size_t)np, hipMemcpyHostToDevice, st));
    if (with_uj) {
        for (int k = 0; k < 3; ++k)
            VPMX_HIP(c, hipMemcpyAsync(f.U + k * f.cap, h + (7 + k) * np, b, hipMemcpyHostToDevice, st));
        for (int k = 0; k < 9; ++k)
            VPMX_HIP(c, hipMemcpyAsync(f.J + k * f.cap, h + (10 + k) * np, b, hipMemcpyHostToDevice, st));
    }
    return sync(c);   // staging is reused: the copies must finish
}

enum { DL_STATE = 1, DL_UJ = 2, DL_SFS = 4 };

int download(vpmx_ctx *c, double *P, const vpmx_layout *L, int what)
{
    DeviceField &f = c->f;
    const int64_t np = f.np;
    if (np == 0) return VPMX_OK;
    VPMX_TRY(ensure_stage(c, (size_t)23 * np * sizeof(double)));
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
        VPMX_HIP(c, hipMemcpyAsync(h + 22 * np, f.C, b, hipMemcpyDeviceToHost, st));
    }
    VPMX_TRY(sync(c));
    const int64_t ld = L->nrows;
    for (int64_t p = 0; p < np; ++p) {
        double *col = P + p * ld;
        if (what & DL_STATE) {
            for (int k = 0; k < 3; ++k) {
                col[L->X + k] = h[k * np + p];
                col[L->Gamma + k] = h[(3 + k) * np + p];
            }
            col[L->sigma] = h[6 * np + p];
        }
        if (what & DL_UJ) {
            for (int k = 0; k < 3; ++k) col[L->U + k] = h[(7 + k) * np + p];
            for (int k = 0; k < 9; ++k) col[L->J + k] = h[(10 + k) * np + p];
        }
        if (sfs) {
            if (L->SFS >= 0)
                for (int k = 0; k < 3; ++k) col[L->SFS + k] = h[(19 + k) * np + p];
            if (L->C >= 0) col[L->C] = h[22 * np + p];
        }
    }
    return VPMX_OK;
}

// ---- device sequences (enqueue only; callers sync) ---------------------------------
int enqueue_uj(vpmx_ctx *c, const vpmx_uj_params *P, int reset)
{
    const DeviceField &f = c->f;
    if (P->precision == VPMX_FP64) {
        VPMX_HIP(c, launch_pack_f64(c->stream, f, c->s));
        TimedScope ts(c, T_UJ);
        VPMX_HIP(c, launch_uj_f64(c->stream, P->kernel, f, c->s, reset));
    } else {
        VPMX_HIP(c, launch_sort_pack_mixed(c->stream, P->kernel, f, c->s));
        TimedScope ts(c, T_UJ);
        VPMX_HIP(c, launch_uj_mixed(c->stream, P->kernel, f, c->s, reset));
    }
    return VPMX_OK;
}

int enqueue_sfs(vpmx_ctx *c, int kernel, int transposed, double Cs, int clip, bool coeff)
{
    {
        TimedScope ts(c, T_ESTR);
        VPMX_HIP(c, launch_estr_f64(c->stream, kernel, transposed, c->f, c->s));
    }
    if (coeff) VPMX_HIP(c, launch_sfs_coeff(c->stream, c->f, Cs, clip));
    return VPMX_OK;
}

// Low-storage RK3 coefficients (Williamson 1980), the scheme of
// docs/src/theory/rvpm.md:364.
constexpr double RK3_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double RK3_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

int enqueue_step(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S, double dt,
                 int relax, bool rk3)
{
    DeviceField &f = c->f;
    const int sfs = S->sfs != VPMX_SFS_NONE;
    if (sfs && U->kernel == VPMX_KERNEL_SINGULAR)
        return fail(c, VPMX_EINVAL, "the SFS model needs a smooth kernel");
    const double zeta0 = zeta0_of(U->kernel);
    const int nst = rk3 ? 3 : 1;
    if (rk3) VPMX_HIP(c, launch_fill(c->stream, f.M, 7 * f.cap, 0.0));
    for (int s = 0; s < nst; ++s) {
        VPMX_TRY(enqueue_uj(c, U, 1));
        if (sfs) VPMX_TRY(enqueue_sfs(c, U->kernel, U->transposed, S->Cs, S->clip_backscatter, true));
        TimedScope ts(c, T_UPDATE);
        VPMX_HIP(c, launch_update(c->stream, f, U->transposed, sfs, S->f, S->g, zeta0, S->Uinf,
                                  rk3 ? RK3_A[s] : 0.0, rk3 ? RK3_B[s] : 1.0, dt));
    }
    if (relax && S->relax_kind != VPMX_RELAX_NONE) {
        if (rk3) VPMX_TRY(enqueue_uj(c, U, 1));
        VPMX_HIP(c, launch_relax(c->stream, f, S->relax_kind, S->rlxf));
    }
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
    free_field(c->f);
    free_scratch(c->s);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
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
    if (P->sfs) VPMX_TRY(enqueue_sfs(c, P->kernel, P->transposed, 0.0, 0, false));
    VPMX_TRY(sync(c));
    return download(c, particles, L, DL_UJ | (P->sfs ? DL_SFS : 0));
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
    VPMX_TRY(enqueue_step(c, U, S, dt, relax, rk3));
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
    VPMX_TRY(enqueue_uj(c, P, P->reset));
    if (P->sfs) VPMX_TRY(enqueue_sfs(c, P->kernel, P->transposed, 0.0, 0, false));
    return sync(c);
}

int vpmx_rk3_resident(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S, double dt,
                      int32_t relax)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(enqueue_step(c, U, S, dt, relax, true));
    return sync(c);
}

int vpmx_euler_resident(vpmx_ctx *c, const vpmx_uj_params *U, const vpmx_step_params *S,
                        double dt, int32_t relax)
{
    VPMX_TRY(entry(c));
    VPMX_TRY(check_uj_params(c, U));
    VPMX_TRY(check_step_params(c, S, dt));
    VPMX_TRY(enqueue_step(c, U, S, dt, relax, false));
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

}  // extern "C"
