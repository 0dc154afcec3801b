/*
 * vpmx.h -- C ABI of the MI355X rVPM particle-interaction library (libvpmx.so).
 *
 * Drop-in boundary for FLOWUnsteady's particle-field hooks.  FLOWUnsteady
 * builds its particle field with function-valued hooks
 *
 *   vpm.ParticleField(max_particles, R; kernel=vpm_kernel, UJ=vpm_UJ,
 *                     SFS=vpm_SFS, integration=vpm_integration,
 *                     transposed=vpm_transposed, relaxation=vpm_relaxation,
 *                     fmm=vpm_fmm)           src/FLOWUnsteady_simulation.jl:239-253
 *
 * and calls them as pfield.UJ(pfield) (src/FLOWUnsteady_simulation.jl:544,
 * src/FLOWUnsteady_processing_force.jl:238) and
 * vpm.nextstep(pfield, dt; relax) -> pfield.integration(pfield, dt; relax)
 * (src/FLOWUnsteady_simulation.jl:358).  A Julia `UJ_gpu(pfield; ...)` and
 * `rungekutta3_gpu(pfield, dt; relax)` bind vpmx_uj / vpmx_rk3_step with one
 * `ccall` each, passing pfield.particles (a column-per-particle Matrix{Float64})
 * and the row offsets of its fields.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a nonzero VPMX_E* code on
 *    failure; vpmx_last_error() gives the message.  Nothing aborts.
 *  - Host pointers are never retained past a call.  The context owns device
 *    buffers, grows them geometrically and reuses them across calls.
 *  - A context is single-threaded: one per host thread, or guard it.
 *  - All calls are blocking unless stated otherwise.
 *  - Host particle matrix: column-major, `nrows` doubles per particle, field
 *    row r of particle p at particles[r + p*nrows] (0-based rows).
 *  - J[k,l] = dU_k/dx_l stored column-major: row J + k + 3*l.
 *  - Results are deterministic: repeated calls are bitwise identical.
 */
#ifndef VPMX_H
#define VPMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define VPMX_OK 0
#define VPMX_EINVAL 1     /* bad argument, layout, sigma <= 0, NaN/Inf input   */
#define VPMX_ENOMEM 2     /* device or pinned-host allocation failed           */
#define VPMX_EHIP 3       /* HIP runtime or kernel-launch error                */
#define VPMX_ENODEV 4     /* no usable GPU / bad device ordinal                */
#define VPMX_EINJECT 5    /* fault injected with vpmx_inject_fault             */
#define VPMX_EUNSUPPORTED 6

/* ---- enumerations ---------------------------------------------------------- */
/* Radial basis / regularising kernels: vpm.singular, vpm.gaussian,
 * vpm.gaussianerf, vpm.winckelmans (src/FLOWUnsteady_simulation.jl:37,128). */
#define VPMX_KERNEL_SINGULAR 0
#define VPMX_KERNEL_GAUSSIAN 1
#define VPMX_KERNEL_GAUSSIANERF 2
#define VPMX_KERNEL_WINCKELMANS 3

/* Arithmetic of the pairwise kernels. */
#define VPMX_FP64 0       /* FP64 throughout (the correctness path)            */
#define VPMX_MIXED 1      /* FP32 pair math on block-relative coordinates,
                             FP32 tile partials folded into FP64 accumulators */

/* Relaxation schemes: vpm.norelaxation, vpm.pedrizzetti,
 * vpm.correctedpedrizzetti (src/FLOWUnsteady_simulation.jl:44,135). */
#define VPMX_RELAX_NONE 0
#define VPMX_RELAX_PEDRIZZETTI 1
#define VPMX_RELAX_CORRECTED 2

/* SFS models (src/FLOWUnsteady_simulation.jl:39,130). */
#define VPMX_SFS_NONE 0
#define VPMX_SFS_CONSTANT 1   /* C = Cs with Estr, optional clipping          */
#define VPMX_SFS_DYNAMIC 2    /* C from the dynamic procedure (exact
                                 core-size derivatives, test-filter level
                                 alpha, Lagrangian average), clipped         */

/* Viscous schemes: vpm.Inviscid, vpm.CoreSpreading, vpm.ParticleStrengthExchange
 * (src/FLOWUnsteady_simulation.jl:42). */
#define VPMX_VISCOUS_NONE 0
#define VPMX_VISCOUS_CORESPREADING 1
#define VPMX_VISCOUS_PSE 2

/* Schedules of the batched pairwise pass (vpmx_batch_set_schedule). */
#define VPMX_SCHED_FLAT 0     /* one workgroup per work item, largest field first  */
#define VPMX_SCHED_QUEUE 1    /* persistent grid draining a device work queue      */

/* ---- types ----------------------------------------------------------------- */
typedef struct vpmx_ctx vpmx_ctx;

/* Row offsets (0-based) of the particle fields inside one particle column;
 * -1 marks a field the caller does not hold.  Filled by the Julia shim from
 * FLOWVPM's *_INDEX constants. */
typedef struct {
    int32_t nrows;        /* leading dimension: doubles per particle column */
    int32_t X;            /* 3 rows: position                              */
    int32_t Gamma;        /* 3 rows: vectorial circulation                 */
    int32_t sigma;        /* 1 row : smoothing radius                      */
    int32_t vol;          /* 1 row : volume (-1 allowed)                   */
    int32_t U;            /* 3 rows: velocity (output)                     */
    int32_t J;            /* 9 rows: velocity Jacobian (output)            */
    int32_t SFS;          /* 3 rows: Estr (output when sfs != 0; -1 allowed) */
    int32_t C;            /* 3 rows: SFS coefficient and the Lagrangian
                             averages of the dynamic procedure's numerator
                             and denominator (-1 allowed)                  */
    int32_t static_;      /* 1 row : static flag, nonzero = static (-1 allowed) */
} vpmx_layout;

/* Settings of one UJ evaluation (the ParticleField's kernel/UJ/fmm fields). */
typedef struct {
    int32_t kernel;       /* VPMX_KERNEL_*                                  */
    int32_t precision;    /* VPMX_FP64 or VPMX_MIXED                        */
    int32_t transposed;   /* stretching scheme used by the SFS pass         */
    int32_t reset;        /* 1: overwrite U,J; 0: accumulate into them      */
    int32_t sfs;          /* VPMX_SFS_*: also evaluate Estr into SFS rows   */
    int32_t fmm;          /* 0 direct all-pairs, 1 fast multipole           */
    int32_t p, ncrit;     /* FMM expansion order (0..12) and leaf size      */
    double theta;         /* FMM multipole acceptance: (R_A+R_B) < theta d  */
    double sigma_cut;     /* FMM: M2L also needs a gap of sigma_cut * max
                             sigma of the source cell.  0: the purely
                             geometric criterion (FLOWVPM's FMM with
                             nonzero_sigma = false, FLOWUnsteady's default,
                             src/FLOWUnsteady_simulation.jl:43); < 0: the
                             kernel's default (gaussian-erf 7, gaussian 3,
                             Winckelmans 20, singular 0)                   */
} vpmx_uj_params;

/* Settings of one time step (formulation, freestream, SFS, relaxation). */
typedef struct {
    double f, g;          /* formulation: rVPM f=0 g=1/5, cVPM f=g=0        */
    double Uinf[3];       /* freestream at the step's start time            */
    int32_t sfs;          /* VPMX_SFS_*                                     */
    double Cs;            /* constant SFS coefficient                       */
    int32_t clip_backscatter; /* clip C to 0 where C Gamma.Estr < 0         */
    int32_t relax_kind;   /* VPMX_RELAX_*                                   */
    double rlxf;          /* relaxation factor                              */
    /* dynamic SFS (sfs = VPMX_SFS_DYNAMIC, docs/src/theory/rvpm.md:264-296):
     * C = <Gamma.L>/<Gamma.m> with exact derivatives D = sigma d/dsigma of
     * every core, Gamma.L = (3 alpha - 2)/sigma Gamma^T D[J] Gamma (the
     * unresolved u - ubar taken at the test filter alpha sigma to first order
     * in 1 - alpha) and Gamma.m = sigma^2/zeta(0) Gamma.D[E]; sfs_rlxf is the
     * Lagrangian-average weight of the newest sample, [minC, maxC] the bounds
     * of C.  With the FMM the derivatives are near-field sums (its far field
     * uses the singular kernel, which has no core). */
    double alpha, sfs_rlxf, minC, maxC;
    /* viscous diffusion nu lap(omega) (vpm_viscous, src/FLOWUnsteady_simulation.jl:42,229-234):
     * VPMX_VISCOUS_PSE adds the particle-strength-exchange rate to dGamma/dt;
     * VPMX_VISCOUS_CORESPREADING adds nu/sigma to dsigma/dt and, after the
     * step, when a free particle's core exceeds beta*sgm0, resets every free
     * core to sgm0 and re-fits the strengths to the vorticity at the particles
     * by conjugate gradients (itmax iterations, relative residual tol). */
    int32_t viscous;      /* VPMX_VISCOUS_*                                 */
    double nu;            /* kinematic viscosity                            */
    double sgm0, beta;    /* core spreading: reset core size and trigger    */
    int32_t itmax;        /* core spreading: CG iterations                  */
    double tol;           /* core spreading: CG relative residual           */
} vpmx_step_params;

/* ---- context ----------------------------------------------------------------- */
/* No reference counterpart: the reference keeps state in Julia objects. */
int vpmx_create(vpmx_ctx **out, int32_t device);
void vpmx_destroy(vpmx_ctx *ctx);
/* Message of the last failing call on ctx (or of vpmx_create when ctx is NULL). */
const char *vpmx_last_error(const vpmx_ctx *ctx);
const char *vpmx_version(void);

/* ---- drop-in entry points (host particle matrix) ----------------------------- */
/* Replaces FLOWVPM's UJ_direct(pfield) / UJ_fmm(pfield) as called at
 * src/FLOWUnsteady_simulation.jl:544 and src/FLOWUnsteady_processing_force.jl:238:
 * U and J of all np particles (static and probes included) induced by all np
 * particles, coincident pairs skipped, source sigma in g, no freestream.
 * When params->sfs is set, also writes Estr into the SFS rows.  U, J (and SFS)
 * rows are valid in `particles` on return. */
int vpmx_uj(vpmx_ctx *ctx, double *particles, int64_t np, const vpmx_layout *layout,
            const vpmx_uj_params *params);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 direct UJ of vpmx_uj (same kernels, same
 * semantics, params->precision = VPMX_FP64, fmm = 0, sfs = 0) on n particles
 * already in device memory, SoA of stride n: X, Gamma (3 x n), sigma (n) in,
 * U (3 x n), J (9 x n; J[(k + 3l) n + i] = dU_k/dx_l of particle i) out,
 * accumulated when params->reset = 0.  It calls no allocator: its scratch is a
 * fixed 48 MiB block of every device's memory, set up by vpmx_create, so
 * n <= VPMX_NOALLOC_MAX_N (larger n fails with VPMX_EINVAL).  Asynchronous:
 * enqueued on `stream` (a hipStream_t; NULL: the context's stream), results
 * valid once that stream is synchronized.  Every call on one device shares the
 * one block, which keeps the last call's contents: calls that overlap on one
 * GPU (other streams, other contexts, other threads) race. */
#define VPMX_NOALLOC_MAX_N 131072
int vpmx_uj_device_noalloc(vpmx_ctx *ctx, int64_t n, const double *X, const double *Gamma,
                           const double *sigma, double *U, double *J,
                           const vpmx_uj_params *params, void *stream);

/* Replaces FLOWVPM's rungekutta3(pfield, dt; relax) as selected by
 * vpm_integration (src/FLOWUnsteady_simulation.jl:40,131,242) and called
 * through vpm.nextstep (src/FLOWUnsteady_simulation.jl:358): three low-storage
 * stages of UJ (+SFS) and a fused FP64 update of non-static particles, then,
 * if relax, a fresh UJ and the relaxation.  Uploads once, keeps the state
 * resident in HBM across the stages, writes X, Gamma, sigma, U, J (and SFS, C)
 * back into `particles`. */
int vpmx_rk3_step(vpmx_ctx *ctx, double *particles, int64_t np, const vpmx_layout *layout,
                  const vpmx_uj_params *uj, const vpmx_step_params *step, double dt,
                  int32_t relax);

/* Replaces FLOWVPM's euler(pfield, dt; relax) (src/FLOWUnsteady_simulation.jl:40). */
int vpmx_euler_step(vpmx_ctx *ctx, double *particles, int64_t np, const vpmx_layout *layout,
                    const vpmx_uj_params *uj, const vpmx_step_params *step, double dt,
                    int32_t relax);

/* ---- resident state (particles kept in HBM between calls) -------------------- */
/* Upload np particles from a host matrix into the context's device state. */
int vpmx_upload(vpmx_ctx *ctx, const double *particles, int64_t np, const vpmx_layout *layout);
/* Write the device state (X, Gamma, sigma, U, J and, when present in the
 * layout, SFS and C) back into a host matrix. */
int vpmx_download(vpmx_ctx *ctx, double *particles, int64_t np, const vpmx_layout *layout);
/* UJ / time steps on the resident state (same semantics as above). */
int vpmx_uj_resident(vpmx_ctx *ctx, const vpmx_uj_params *params);
int vpmx_rk3_resident(vpmx_ctx *ctx, const vpmx_uj_params *uj, const vpmx_step_params *step,
                      double dt, int32_t relax);
int vpmx_euler_resident(vpmx_ctx *ctx, const vpmx_uj_params *uj, const vpmx_step_params *step,
                        double dt, int32_t relax);
/* U (m,3) and J (m,9) (row-major per target, J[k + 3l] = dU_k/dx_l) induced
 * by the resident particles at m arbitrary points, without adding them to the
 * field: the probe evaluation of Vvpm_on_Xs (src/FLOWUnsteady_simulation.jl:494-570),
 * which otherwise appends zero-strength probes and evaluates all (N+m)^2
 * pairs.  Direct scheme only; coincident points are skipped. */
int vpmx_uj_targets(vpmx_ctx *ctx, const double *targets, int64_t m, double *U, double *J,
                    const vpmx_uj_params *params);
/* Enstrophy 0.5 sum_p Gamma_p . omega(x_p), omega = curl U from the resident J
 * (vpm.monitor_enstrophy, src/FLOWUnsteady_monitors.jl:573-574,614), and the
 * energy sum_p U_p . (x_p x Gamma_p). */
int vpmx_diagnostics(vpmx_ctx *ctx, double *enstrophy, double *energy);
/* Number of particles currently resident. */
int64_t vpmx_resident_np(const vpmx_ctx *ctx);

/* Free-vortex force of bound vortices on the resident field, FLOWUnsteady's
 * _Ffv_direct(source, target) (src/FLOWUnsteady_processing_force.jl:860-929)
 * with the bound particles as `source` and the resident particles (static
 * included) as `target`: for bound particle b of `bound` (nb columns in
 * `layout`), with dX = x_f - x_b and g = g(r / sigma_b),
 *   out[6 b + 0..2] = Ffv/rho  = sum_f U_b(x_f) x Gamma_f,
 *   out[6 b + 3..5] = Fbxf/rho = sum_f g/(4 pi r^3) dX x (Gamma_b x Gamma_f),
 * (the values the reference stores in P.M[1:6]); coincident pairs skipped.
 * FP64, fixed summation order.  kernel: VPMX_KERNEL_*. */
int vpmx_ffv(vpmx_ctx *ctx, const double *bound, int64_t nb, const vpmx_layout *layout,
             int32_t kernel, double *out);

/* Core spreading's spatial adaptation on the resident field (what
 * vpm.rbf_conjugategradient provides to vpm.CoreSpreading): when force != 0
 * or a free particle's core exceeds step->beta * step->sgm0, every free
 * (non-static) core is set to sgm0 and the free strengths are re-fitted so
 * that sum_j Gamma'_j zeta_sgm0(x_i - x_j) equals the free particles' vorticity
 * sum_j Gamma_j zeta_sigma_j(x_i - x_j) at every free particle, by conjugate
 * gradients from the current strengths (step->itmax, step->tol).  FP64 sums:
 * all pairs when uj->fmm = 0 (vpm.zeta_direct); with uj->fmm = 1 the near
 * field of an FMM tree built at the current positions with uj's ncrit, theta
 * and sigma_cut (vpm.zeta_fmm, the setting of examples/rotorhover/rotorhover.jl:170),
 * whose far field of zeta is neglected.  The same near lists carry the PSE
 * sums of an FMM time step.  The kernel must be gaussian-erf.  *did (optional) = 1 when the
 * reset ran, *iters the CG iterations, res[3] (optional) the final relative
 * residual per component. */
int vpmx_core_spreading_reset(vpmx_ctx *ctx, const vpmx_uj_params *uj,
                              const vpmx_step_params *step, int32_t force, int32_t *did,
                              int32_t *iters, double *res);

/* ---- batched fields ------------------------------------------------------------- */
/* Many independent particle fields of uneven size (a sweep of rotor or wing
 * cases, each of which run_simulation advances with its own nextstep,
 * src/FLOWUnsteady_simulation.jl:339-365) resident as one concatenation and
 * advanced together: one pairwise launch covers every field, and no field
 * ever sees another's particles.  particles[k] is field k's host matrix with
 * np[k] columns, all in the same layout.  A batch of one field is a single
 * resident field.  The direct scheme only; the SFS model and the FMM are
 * rejected with VPMX_EUNSUPPORTED for more than one field.  The resident entry
 * points above (vpmx_uj_resident, vpmx_rk3_resident, vpmx_euler_resident,
 * vpmx_diagnostics, which then sums over the fields) accept a resident batch. */
int vpmx_batch_upload(vpmx_ctx *ctx, int32_t nfields, const double *const *particles,
                      const int64_t *np, const vpmx_layout *layout);
int vpmx_batch_download(vpmx_ctx *ctx, int32_t nfields, double *const *particles,
                        const int64_t *np, const vpmx_layout *layout);
/* UJ of every field: upload, one evaluation, U and J written back (vpmx_uj per field). */
int vpmx_batch_uj(vpmx_ctx *ctx, int32_t nfields, double *const *particles, const int64_t *np,
                  const vpmx_layout *layout, const vpmx_uj_params *params);
/* One RK3 step of every field (vpmx_rk3_step per field).  Uinf: 3 values per
 * field (the freestream of each case), or NULL for step->Uinf everywhere. */
int vpmx_batch_rk3_step(vpmx_ctx *ctx, int32_t nfields, double *const *particles,
                        const int64_t *np, const vpmx_layout *layout,
                        const vpmx_uj_params *uj, const vpmx_step_params *step,
                        const double *Uinf, double dt, int32_t relax);
/* The same step on the resident batch. */
int vpmx_batch_rk3_resident(vpmx_ctx *ctx, const vpmx_uj_params *uj,
                            const vpmx_step_params *step, const double *Uinf, double dt,
                            int32_t relax);
/* Number of resident fields (0 when nothing was uploaded). */
int32_t vpmx_batch_nfields(const vpmx_ctx *ctx);
/* Schedule of the batched pairwise pass: VPMX_SCHED_FLAT or VPMX_SCHED_QUEUE,
 * workgroups of block = 64, 128 or 256 targets, and for the queue the grid
 * (0 = resident occupancy x compute units). */
int vpmx_batch_set_schedule(vpmx_ctx *ctx, int32_t schedule, int32_t block, int32_t grid);
int vpmx_batch_get_schedule(vpmx_ctx *ctx, int32_t *schedule, int32_t *block, int32_t *grid);
/* Autotuner: times the pairwise pass of the resident batch under every
 * candidate schedule (flat and queue at block 64, 128, 256; flat only for
 * FP64), best of `reps` runs each, keeps the fastest.  Writes the candidates'
 * milliseconds into ms[0 .. min(nms, ncand) - 1] in that order (flat 64, queue 64,
 * flat 128, queue 128, flat 256, queue 256; FP64: flat 64, 128, 256) and
 * their count into *ncand.  Leaves U and J of the resident batch evaluated. */
int vpmx_batch_autotune(vpmx_ctx *ctx, const vpmx_uj_params *params, int32_t reps, double *ms,
                        int32_t nms, int32_t *ncand);

/* ---- graphs ---------------------------------------------------------------------- */
/* on != 0: time steps on the direct scheme (vpmx_rk3_*, vpmx_euler_*,
 * vpmx_batch_rk3_*) run their device work -- every substep's UJ, SFS and fused
 * update and the relaxation -- as one instantiated hipGraph, captured on first
 * use and replayed while the settings, dt, particle count and buffers stay the
 * same (anything else re-captures).  The freestream and the Morton order are
 * set before each replay.  Results are bitwise those of the eager path. */
int vpmx_set_graph(vpmx_ctx *ctx, int32_t on);

/* ---- instrumentation ----------------------------------------------------------- */
/* Time every launch of the pairwise kernels with HIP events on the context's
 * stream.  which: 0 = UJ pair kernel, 1 = Estr pair kernel, 2 = fused update,
 * 3 = FMM device evaluation (P2M through P2P), 4 = graph replay of a whole step,
 * 5 = the dynamic SFS's core-size derivatives D[U], D[J]
 * (the kernels inside a graph are not timed one by one).
 * Returns the summed milliseconds and the launch count since the last reset. */
int vpmx_set_timing(vpmx_ctx *ctx, int32_t on);
int vpmx_get_timing(vpmx_ctx *ctx, int32_t which, double *total_ms, int64_t *count);
/* Statistics of the last FMM evaluation on ctx, in this order: tree build ms
 * (on the device, wall time including its one count read-back per level),
 * traversal ms (on the device, including the list sort), host-to-device tree
 * copy ms (0: the tree is built where it is used), cells, leaves, M2L pairs,
 * P2P leaf pairs, P2P particle pairs, device evaluation ms (when timing is on).
 * Writes min(n, 9) values. */
int vpmx_fmm_stats(vpmx_ctx *ctx, double *out, int32_t n);
/* Host-only: build the FMM tree and traversal lists for positions X (n,3)
 * and sigma exactly as vpmx_uj does with fmm = 1, and report cells, leaves,
 * M2L pairs, P2P leaf pairs and P2P particle pairs in out[0..4] (for checking
 * the plan without a GPU). */
int vpmx_fmm_plan_counts(int64_t n, const double *X, const double *sigma, int32_t ncrit,
                         double theta, double sigma_cut, int64_t *out);
/* The context's HIP stream (hipStream_t), for callers that order work on it. */
void *vpmx_stream(vpmx_ctx *ctx);
/* Make the next `count` entry-point calls fail with VPMX_EINJECT before any
 * device work (error-path testing). */
int vpmx_inject_fault(vpmx_ctx *ctx, int32_t count);
/* Host evaluation of the FP32 radial coefficients A = g/r^3 and
 * B = g'/(sigma r^4) - 3 g/r^5 exactly as the mixed-precision kernel's
 * near-field branch computes them (for unit tests of their accuracy; the
 * hardware exp/rcp/rsqrt are replaced by the C library's). */
int vpmx_radial_f32(int32_t kernel, float r, float sigma, float *A, float *B);

#ifdef __cplusplus
}
#endif

#endif /* VPMX_H */
