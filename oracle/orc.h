/*
 * orc.h -- declarations of the CPU FP64 oracle (vpm_oracle.c, fmm_oracle.c,
 * viscous_oracle.c).
 *
 * TEST INFRASTRUCTURE ONLY: the checker, loaded by tests/, smoke() and the
 * benchmark's CPU-baseline leg; the product never includes or links it.
 * Arrays are AoS per particle: X[3*i+k], J[9*i+k+3*l] with J[k,l] = dU_k/dx_l.
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>

struct orc_fmm_stats;   /* fmm_oracle.c: timings and counts of one evaluation */

typedef struct {
    int64_t np;
    double *X, *G, *sigma, *U, *J, *E, *M;    /* M: 7 registers per particle */
    double *C;                                /* 3 per particle: C, <num>, <den> */
    const uint8_t *is_static;                  /* may be NULL: no static particles */
    const double *vol;                         /* may be NULL: PSE needs it */
} orc_field;

typedef struct {
    int kernel, transposed;
    double f, g;               /* formulation: rVPM f=0 g=1/5, cVPM f=g=0 */
    double Uinf[3];
    int sfs;                   /* 0 none, 1 constant Cs, 2 dynamic; Estr_direct */
    double Cs;
    int clip_backscatter;
    int relax_kind;            /* 0 none, 1 pedrizzetti, 2 corrected */
    double rlxf;
    double alpha, sfs_rlxf, minC, maxC;   /* dynamic procedure */
    int viscous;               /* 0 inviscid, 1 core spreading, 2 PSE */
    double nu, sgm0, beta;     /* viscosity; core spreading reset size and trigger */
    int itmax;                 /* core spreading: CG iterations */
    double tol;                /* core spreading: CG relative residual */
    /* UJ scheme: 0 UJ_direct, 1 UJ_fmm (orc_uj_fmm with p, ncrit, theta and the
     * acceptance gap sigcut; Estr and the core-size derivatives then run over
     * the FMM's near-field lists, orc_fmm_near) */
    int fmm, p, ncrit;
    double theta, sigcut;
} orc_params;

/* fmm_oracle.c */
int orc_uj_fmm(int kernel, int p, int ncrit, double theta, double sigcut, int64_t n,
               const double *X, const double *G, const double *sig, double *U, double *J,
               const int64_t *sample, int64_t nsample, struct orc_fmm_stats *stats);
int orc_fmm_near(int kernel, int transposed, int what, int ncrit, double theta, double sigcut,
                 int64_t n, const double *X, const double *G, const double *sig, const double *J,
                 const double *DJ, double *out1, double *out2);

/* viscous_oracle.c */
void orc_pse(int kernel, double nu, int64_t n, const double *X, const double *G,
             const double *sig, const double *vol, int64_t nt, const int64_t *tidx, double *dG);
void orc_zeta_sum(int kernel, int64_t nt, const double *tX, int64_t ns, const double *sX,
                  const double *sG, const double *ssig, double s0, double *w);
int orc_rbf_cg(int kernel, int64_t n, const double *X, double s0, const double *w, double *G,
               int itmax, double tol, double *res);

/* vpm_oracle.c */
void orc_kernel(int kernel, double rho, double *zeta, double *g, double *dg);
double orc_dzeta(int kernel, double rho);
void orc_kernel_dsig(int kernel, double rho, double *h, double *dh, double *zD);
double orc_zeta0(int kernel);
void orc_pair_uj(int kernel, const double *xt, const double *xs, const double *gs, double sgm,
                 double *U, double *J);
void orc_pair_dj(int kernel, const double *xt, const double *xs, const double *gs, double sgm,
                 double *DU, double *DJ);
void orc_uj_direct(int kernel, int64_t nt, const double *tX, int64_t ns, const double *sX,
                   const double *sG, const double *ssig, double *U, double *J);
void orc_dj_direct(int kernel, int64_t nt, const double *tX, int64_t ns, const double *sX,
                   const double *sG, const double *ssig, double *DU, double *DJ);
void orc_ffv_direct(int kernel, int64_t nb, const double *bX, const double *bG,
                    const double *bsig, int64_t nf, const double *fX, const double *fG,
                    double *out);
void orc_stretching(int transposed, const double *J, const double *G, double *S);
void orc_pair_estr(int kernel, int transposed, const double *xt, const double *tJ,
                   const double *tDJ, const double *xs, const double *gs, double sg,
                   const double *sJ, const double *sDJ, double *e, double *de);
void orc_estr_direct(int kernel, int transposed, int64_t nt, const double *tX, const double *tJ,
                     int64_t ns, const double *sX, const double *sG, const double *ssig,
                     const double *sJ, double *E);
void orc_estr_dyn_direct(int kernel, int transposed, int64_t nt, const double *tX,
                         const double *tJ, const double *tDJ, int64_t ns, const double *sX,
                         const double *sG, const double *ssig, const double *sJ,
                         const double *sDJ, double *E, double *DE);
void orc_rhs(int transposed, double f, double gcoef, double zeta0, const double *Uinf,
             const double *U, const double *J, const double *G, double sigma, double C,
             const double *E, double *dX, double *dG, double *dsigma);
void orc_lsrk3_linear(double lre, double lim, double dt, double *yre, double *yim);
void orc_relax(int kind, double rlxf, const double *J, double *G);
int orc_core_spreading_reset(const orc_params *P, orc_field *F, int *iters);
void orc_dyn_terms(double alpha, double zeta0, double sigma, const double *G, const double *DJ,
                   const double *DE, double *num, double *den);
void orc_rk3_step(const orc_params *P, orc_field *F, double dt, int relax);
void orc_euler_step(const orc_params *P, orc_field *F, double dt, int relax);
double orc_enstrophy(int64_t np, const double *G, const double *J);
double orc_energy(int64_t np, const double *X, const double *G, const double *U);
int orc_num_threads(void);

/* fmm_oracle.c */
int orc_fmm_deriv(int P, const double r[3], double *b, int *mi_out);
struct orc_near_plan;
struct orc_near_plan *orc_near_plan_build(int ncrit, double theta, double sigcut, int64_t n,
                                          const double *X, const double *sig);
void orc_near_plan_free(struct orc_near_plan *pl);
void orc_near_zeta(const struct orc_near_plan *pl, int kernel, const double *X, const double *G,
                   const double *ssig, double s0, const uint8_t *mask, double *w);
void orc_near_pse(const struct orc_near_plan *pl, int kernel, double nu, const double *X,
                  const double *G, const double *sig, const double *vol, double *V);

/* viscous_oracle.c */
double orc_eta(int kernel, double rho);

#endif /* ORC_H */
