/*
 * vpm_oracle.c -- CPU FP64 restatement of the rVPM particle-interaction step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  The
 * product (flowunsteady_amd/, libvpmx.so) never links or calls it.
 *
 * The path it restates is FLOWVPM's UJ_direct / Estr_direct / rungekutta3 /
 * euler / relaxation, which FLOWUnsteady drives through
 *   vpm.ParticleField(...; kernel, UJ, SFS, integration, transposed,
 *                     relaxation)            src/FLOWUnsteady_simulation.jl:239-253
 *   vpm.nextstep(pfield, dt; relax)          src/FLOWUnsteady_simulation.jl:358
 *   pfield.UJ(pfield)                        src/FLOWUnsteady_simulation.jl:544
 * FLOWVPM itself is not vendored (Project.toml:25 pins FLOWVPM = "4.0"), so
 * every formula below follows the math the reference documents in
 * docs/src/theory/rvpm.md and the pair-loop form of
 * src/FLOWUnsteady_processing_force.jl:879-929.
 *
 * Layout: particle arrays are AoS per particle, e.g. X[3*i+k], J[9*i+k+3*l]
 * with J[k,l] = dU_k/dx_l (column-major 3x3, as a Julia 3x3 matrix).
 *
 * Pins (each checked by tests/test_oracle.py):
 *   orc_kernel        -- closed forms vs quadrature g(rho)=int_0^rho 4 pi s^2 zeta ds
 *                        (tests/golden/kernels.json), g(0)=0, g(inf)=1.
 *   orc_uj_direct     -- vs an independent 50-digit vector-potential route
 *                        u = curl sum_j Gamma_j G_sigma(x-x_j) with mpmath
 *                        numerical derivatives (tests/golden/uj_*.json); J vs
 *                        central differences of U; singular ring on-axis closed
 *                        form; trace(J)=0.
 *   orc_estr_direct   -- vs golden Estr built from the golden J; transposed vs
 *                        classic convention vs finite differences of U.
 *   orc_rhs           -- |Gamma| sigma^2 invariant of rVPM (f=0, g=1/5);
 *                        classic/transposed stretching vs finite differences.
 *   orc_lsrk3_linear  -- one step of y'=lambda*y equals 1+z+z^2/2+z^3/6.
 *   orc_rk3_step      -- third-order convergence on a vortex ring; uniform
 *                        free-stream translation is exact.
 *   orc_relax         -- rlxf=0 identity, rlxf=1 gives |Gamma| omega_hat,
 *                        corrected variant preserves |Gamma|, omega vs curl of U
 *                        by finite differences.
 *   orc_enstrophy     -- 0.5 sum Gamma.omega (src/FLOWUnsteady_monitors.jl:573-574).
 *   orc_ffv_direct    -- Ffv part vs sum_f U_b(x_f) x Gamma_f with U_b from
 *                        orc_uj_direct of the single source b; the Gamma_b x
 *                        Gamma_f part vs a numpy restatement; both vs the
 *                        antisymmetry of swapping the roles of b and f.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc.h"

/* 1/(4 pi) exactly as the reference pair loop writes it
 * (src/FLOWUnsteady_processing_force.jl:900-902). */
#define ORC_CONST4 0.07957747154594767
#define ORC_PI 3.14159265358979323846

enum { ORC_SINGULAR = 0, ORC_GAUSSIAN = 1, ORC_GAUSSIANERF = 2, ORC_WINCKELMANS = 3 };

/* Radial basis zeta(rho), regularising function g(rho) and g'(rho).
 * docs/src/theory/rvpm.md:84 (zeta_sigma = sigma^-3 zeta(r/sigma)),
 * rvpm.md:89-97 (g_sigma in the regularised Biot-Savart law),
 * rvpm.md:365 (Gaussian filter), rvpm.md:380 (Winckelmans & Leonard 1993).
 * All satisfy g'(rho) = 4 pi rho^2 zeta(rho). */
void orc_kernel(int kernel, double rho, double *zeta, double *g, double *dg)
{
    double z = 0.0, gg = 1.0, d = 0.0;
    switch (kernel) {
    case ORC_SINGULAR:
        z = 0.0; gg = 1.0; d = 0.0;
        break;
    case ORC_GAUSSIAN: {
        double e = exp(-rho * rho * rho);
        z = 3.0 / (4.0 * ORC_PI) * e;
        gg = 1.0 - e;
        d = 3.0 * rho * rho * e;
        break;
    }
    case ORC_GAUSSIANERF: {
        const double c2 = 0.7978845608028654;   /* sqrt(2/pi) */
        double e = exp(-rho * rho / 2.0);
        z = 1.0 / pow(2.0 * ORC_PI, 1.5) * e;
        gg = erf(rho / 1.4142135623730951) - c2 * rho * e;
        d = c2 * rho * rho * e;
        break;
    }
    case ORC_WINCKELMANS: {
        double q = rho * rho + 1.0;
        double q52 = pow(q, 2.5), q72 = pow(q, 3.5);
        z = 15.0 / (8.0 * ORC_PI) / q72;
        gg = rho * rho * rho * (rho * rho + 2.5) / q52;
        d = 7.5 * rho * rho / q72;
        break;
    }
    default:
        z = NAN; gg = NAN; d = NAN;
    }
    if (zeta) *zeta = z;
    if (g) *g = gg;
    if (dg) *dg = d;
}

/* zeta'(rho), the derivative of the closed forms above:
 *   gaussianerf  -rho zeta,   gaussian  -3 rho^2 zeta,
 *   winckelmans  -7 rho / (rho^2 + 1) zeta,   singular  0. */
double orc_dzeta(int kernel, double rho)
{
    double z;
    orc_kernel(kernel, rho, &z, NULL, NULL);
    switch (kernel) {
    case ORC_GAUSSIAN: return -3.0 * rho * rho * z;
    case ORC_GAUSSIANERF: return -rho * z;
    case ORC_WINCKELMANS: return -7.0 * rho / (rho * rho + 1.0) * z;
    default: return 0.0;
    }
}

/* Derivatives with respect to the core size, every core scaled together
 * (sigma -> lambda sigma, D = d/dlambda at lambda = 1 = sigma d/dsigma), as the
 * dynamic SFS procedure needs them (docs/src/theory/rvpm.md:276-290):
 *   D g(r/sigma)              = h(rho)  = -rho g'(rho) = -4 pi rho^3 zeta(rho)
 *   d h / d rho               = h'(rho) = -4 pi (3 rho^2 zeta + rho^3 zeta')
 *   D [zeta(r/sigma)/sigma^3] = zD(rho) / sigma^3,  zD = -(3 zeta + rho zeta').
 * U and J carry g only through g(r/sigma)/r^3, so D[U] and D[J] are the pair
 * algebra of orc_pair_uj with (g, g') replaced by (h, h'). */
void orc_kernel_dsig(int kernel, double rho, double *h, double *dh, double *zD)
{
    double z;
    orc_kernel(kernel, rho, &z, NULL, NULL);
    const double dz = orc_dzeta(kernel, rho);
    if (h) *h = -4.0 * ORC_PI * rho * rho * rho * z;
    if (dh) *dh = -4.0 * ORC_PI * (3.0 * rho * rho * z + rho * rho * rho * dz);
    if (zD) *zD = -(3.0 * z + rho * dz);
}

double orc_zeta0(int kernel)
{
    double z;
    if (kernel == ORC_SINGULAR) return 1.0;   /* unused: SFS needs a smooth kernel */
    orc_kernel(kernel, 0.0, &z, NULL, NULL);
    return z;
}

/* Pair algebra shared by U, J and their sigma-derivatives: a radial function
 * f(rho) in place of g with its derivative df, so the pair adds
 *   U += -f/(4 pi r^3) dX x Gamma,
 *   J += (df/(sigma r) - 3 f / r^2) * crss_k * dX_l   (radial part)
 *      + (-f/(4 pi r^3)) * eps_{klm} Gamma_m          (Kronecker part)
 * with crss = -(1/(4 pi r^3)) dX x Gamma.  r > 0. */
static void pair_accum(double f, double df, double sgm, double dX1, double dX2, double dX3,
                       double r2, const double *gs, double *U, double *J)
{
    double r = sqrt(r2);
    double r3 = r * r2;
    double c1 = -ORC_CONST4 / r3 * (dX2 * gs[2] - dX3 * gs[1]);
    double c2 = -ORC_CONST4 / r3 * (dX3 * gs[0] - dX1 * gs[2]);
    double c3 = -ORC_CONST4 / r3 * (dX1 * gs[1] - dX2 * gs[0]);
    U[0] += f * c1;
    U[1] += f * c2;
    U[2] += f * c3;
    double aux = df / (sgm * r) - 3.0 * f / r2;
    /* column l = derivative direction */
    J[0] += aux * c1 * dX1; J[1] += aux * c2 * dX1; J[2] += aux * c3 * dX1;
    J[3] += aux * c1 * dX2; J[4] += aux * c2 * dX2; J[5] += aux * c3 * dX2;
    J[6] += aux * c1 * dX3; J[7] += aux * c2 * dX3; J[8] += aux * c3 * dX3;
    double a2 = -ORC_CONST4 * f / r3;
    /* eps_{klm} Gamma_m: (k,l) = (2,1):-G3 (3,1):+G2 (1,2):+G3 (3,2):-G1 (1,3):-G2 (2,3):+G1 */
    J[1] -= a2 * gs[2];
    J[2] += a2 * gs[1];
    J[3] += a2 * gs[2];
    J[5] -= a2 * gs[0];
    J[6] -= a2 * gs[1];
    J[7] += a2 * gs[0];
}

/* One source j acting on one target: accumulates U[3] and J[9].
 * Pair algebra of src/FLOWUnsteady_processing_force.jl:888-905 (dX = x_t - x_s,
 * skip r == 0, source sigma in g, 1/(4 pi) constant), extended to
 * J[k,l] = dU_k/dx_l by the product rule (pair_accum with f = g). */
void orc_pair_uj(int kernel, const double *xt, const double *xs, const double *gs,
                           double sgm, double *U, double *J)
{
    double dX1 = xt[0] - xs[0], dX2 = xt[1] - xs[1], dX3 = xt[2] - xs[2];
    double r2 = dX1 * dX1 + dX2 * dX2 + dX3 * dX3;
    if (r2 == 0.0) return;
    double g, dg;
    orc_kernel(kernel, sqrt(r2) / sgm, NULL, &g, &dg);
    pair_accum(g, dg, sgm, dX1, dX2, dX3, r2, gs, U, J);
}

/* The same pair's core-size derivatives D[U], D[J] (D = sigma d/dsigma with
 * every core scaled together; pair_accum with f = h, df = h' of
 * orc_kernel_dsig).  Coincident pairs skipped, as for U and J. */
void orc_pair_dj(int kernel, const double *xt, const double *xs, const double *gs, double sgm,
                 double *DU, double *DJ)
{
    double dX1 = xt[0] - xs[0], dX2 = xt[1] - xs[1], dX3 = xt[2] - xs[2];
    double r2 = dX1 * dX1 + dX2 * dX2 + dX3 * dX3;
    if (r2 == 0.0) return;
    double h, dh;
    orc_kernel_dsig(kernel, sqrt(r2) / sgm, &h, &dh, NULL);
    pair_accum(h, dh, sgm, dX1, dX2, dX3, r2, gs, DU, DJ);
}

/* UJ_direct: U and J at every target from every source (accumulating; the
 * caller zeroes first, as _reset_particles does at
 * src/FLOWUnsteady_processing_force.jl:237).  Sources and targets may be the
 * same arrays; coincident pairs are skipped. */
void orc_uj_direct(int kernel, int64_t nt, const double *tX, int64_t ns, const double *sX,
                   const double *sG, const double *ssig, double *U, double *J)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; i++) {
        double u[3] = {0, 0, 0}, jj[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t j = 0; j < ns; j++)
            orc_pair_uj(kernel, tX + 3 * i, sX + 3 * j, sG + 3 * j, ssig[j], u, jj);
        for (int k = 0; k < 3; k++) U[3 * i + k] += u[k];
        for (int k = 0; k < 9; k++) J[9 * i + k] += jj[k];
    }
}

/* D[U] and D[J] (core-size derivatives, orc_pair_dj) at every target from
 * every source, accumulating like orc_uj_direct. */
void orc_dj_direct(int kernel, int64_t nt, const double *tX, int64_t ns, const double *sX,
                   const double *sG, const double *ssig, double *DU, double *DJ)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; i++) {
        double u[3] = {0, 0, 0}, jj[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t j = 0; j < ns; j++)
            orc_pair_dj(kernel, tX + 3 * i, sX + 3 * j, sG + 3 * j, ssig[j], u, jj);
        for (int k = 0; k < 3; k++) DU[3 * i + k] += u[k];
        for (int k = 0; k < 9; k++) DJ[9 * i + k] += jj[k];
    }
}

/* Free-vortex force of bound vortices on the free field, the loop of
 * _Ffv_direct (src/FLOWUnsteady_processing_force.jl:879-929): for each bound
 * particle b (sources) and every free particle f (targets, coincident skipped),
 * with dX = x_f - x_b and g from r / sigma_b,
 *   U_b(x_f) = -g/(4 pi r^3) dX x Gamma_b,   out[6b+0..2] += U_b(x_f) x Gamma_f,
 *   out[6b+3..5] += g/(4 pi r^3) dX x (Gamma_b x Gamma_f).
 * out (nb x 6) is overwritten (the reference flushes P.M first). */
void orc_ffv_direct(int kernel, int64_t nb, const double *bX, const double *bG,
                    const double *bsig, int64_t nf, const double *fX, const double *fG,
                    double *out)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < nb; b++) {
        double m[6] = {0, 0, 0, 0, 0, 0};
        const double *xb = bX + 3 * b, *gb = bG + 3 * b;
        for (int64_t f = 0; f < nf; f++) {
            const double *xf = fX + 3 * f, *gf = fG + 3 * f;
            const double d1 = xf[0] - xb[0], d2 = xf[1] - xb[1], d3 = xf[2] - xb[2];
            const double r = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
            if (r == 0.0) continue;
            double zeta, g, dg;
            orc_kernel(kernel, r / bsig[b], &zeta, &g, &dg);
            const double k = g * ORC_CONST4 / (r * r * r);
            const double u1 = -k * (d2 * gb[2] - d3 * gb[1]);
            const double u2 = -k * (d3 * gb[0] - d1 * gb[2]);
            const double u3 = -k * (d1 * gb[1] - d2 * gb[0]);
            m[0] += u2 * gf[2] - u3 * gf[1];
            m[1] += u3 * gf[0] - u1 * gf[2];
            m[2] += u1 * gf[1] - u2 * gf[0];
            const double c1 = gb[1] * gf[2] - gb[2] * gf[1];
            const double c2 = gb[2] * gf[0] - gb[0] * gf[2];
            const double c3 = gb[0] * gf[1] - gb[1] * gf[0];
            m[3] += k * (d2 * c3 - d3 * c2);
            m[4] += k * (d3 * c1 - d1 * c3);
            m[5] += k * (d1 * c2 - d2 * c1);
        }
        for (int k = 0; k < 6; k++) out[6 * b + k] = m[k];
    }
}

/* Vortex stretching from J (docs/src/theory/rvpm.md:105-178 and :363):
 *   classic     S_k = sum_l J[k,l] Gamma_l   ((Gamma . grad) u)
 *   transposed  S_k = sum_l J[l,k] Gamma_l   ((Gamma . grad^T) u), the default
 *               (vpm_transposed = true, src/FLOWUnsteady_simulation.jl:41). */
void orc_stretching(int transposed, const double *J, const double *G, double *S)
{
    for (int k = 0; k < 3; k++) {
        double s = 0.0;
        for (int l = 0; l < 3; l++)
            s += (transposed ? J[l + 3 * k] : J[k + 3 * l]) * G[l];
        S[k] = s;
    }
}

/* One source j of the SFS model of vortex stretching at one target
 * (docs/src/theory/rvpm.md:251-262):
 *   e  += zeta_sigma_j(x_i - x_j) S(J_i - J_j, Gamma_j),
 * and, when the core-size derivatives DJ are given (tDJ and sDJ not NULL),
 *   de += D[zeta_sigma_j] S(J_i - J_j, Gamma_j) + zeta_sigma_j S(DJ_i - DJ_j, Gamma_j),
 * the derivative D = sigma d/dsigma of e (orc_kernel_dsig's zD).  S is the
 * stretching of orc_stretching (transposed or classic). */
void orc_pair_estr(int kernel, int transposed, const double *xt, const double *tJ,
                   const double *tDJ, const double *xs, const double *gs, double sg,
                   const double *sJ, const double *sDJ, double *e, double *de)
{
    double d1 = xt[0] - xs[0], d2 = xt[1] - xs[1], d3 = xt[2] - xs[2];
    double r = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
    double z;
    orc_kernel(kernel, r / sg, &z, NULL, NULL);
    double is3 = 1.0 / (sg * sg * sg);
    double dJ[9], S[3];
    for (int k = 0; k < 9; k++) dJ[k] = tJ[k] - sJ[k];
    orc_stretching(transposed, dJ, gs, S);
    for (int k = 0; k < 3; k++) e[k] += z * is3 * S[k];
    if (tDJ && sDJ) {
        double zD, dS[3];
        orc_kernel_dsig(kernel, r / sg, NULL, NULL, &zD);
        for (int k = 0; k < 9; k++) dJ[k] = tDJ[k] - sDJ[k];
        orc_stretching(transposed, dJ, gs, dS);
        for (int k = 0; k < 3; k++) de[k] += zD * is3 * S[k] + z * is3 * dS[k];
    }
}

/* Estr_direct: SFS model of vortex stretching (docs/src/theory/rvpm.md:251-262)
 *   E(x_i) = sum_j zeta_sigma_j(x_i - x_j) (Gamma_j . grad)(U(x_i) - U(x_j)),
 * transposed or classic as above.  Accumulates into E. */
void orc_estr_direct(int kernel, int transposed, int64_t nt, const double *tX, const double *tJ,
                     int64_t ns, const double *sX, const double *sG, const double *ssig,
                     const double *sJ, double *E)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; i++) {
        double e[3] = {0, 0, 0};
        for (int64_t j = 0; j < ns; j++)
            orc_pair_estr(kernel, transposed, tX + 3 * i, tJ + 9 * i, NULL, sX + 3 * j,
                          sG + 3 * j, ssig[j], sJ + 9 * j, NULL, e, NULL);
        for (int k = 0; k < 3; k++) E[3 * i + k] += e[k];
    }
}

/* Estr and its core-size derivative D[E] from J and D[J] (orc_pair_estr),
 * accumulating into E and DE. */
void orc_estr_dyn_direct(int kernel, int transposed, int64_t nt, const double *tX,
                         const double *tJ, const double *tDJ, int64_t ns, const double *sX,
                         const double *sG, const double *ssig, const double *sJ,
                         const double *sDJ, double *E, double *DE)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; i++) {
        double e[3] = {0, 0, 0}, de[3] = {0, 0, 0};
        for (int64_t j = 0; j < ns; j++)
            orc_pair_estr(kernel, transposed, tX + 3 * i, tJ + 9 * i, tDJ + 9 * i, sX + 3 * j,
                          sG + 3 * j, ssig[j], sJ + 9 * j, sDJ + 9 * j, e, de);
        for (int k = 0; k < 3; k++) {
            E[3 * i + k] += e[k];
            DE[3 * i + k] += de[k];
        }
    }
}

/* Right-hand side of the rVPM governing equations for one particle
 * (docs/src/theory/rvpm.md:107-178 with the general f, g of :196-236):
 *   Z       = [ (g+f)/(1+3f) S.Gamma - f/(1+3f) (C sigma^3/zeta0) E.Gamma ] / |Gamma|^2
 *   dX/dt   = U + Uinf
 *   dG/dt   = S - 3 Z Gamma - (C sigma^3/zeta0) E
 *   dsg/dt  = -sigma Z
 * rVPM is f = 0, g = 1/5; classic VPM is f = g = 0.  A particle with
 * |Gamma| = 0 gets Z = 0 (probes carry Gamma = 0, simulation.jl:572). */
void orc_rhs(int transposed, double f, double gcoef, double zeta0, const double *Uinf,
             const double *U, const double *J, const double *G, double sigma, double C,
             const double *E, double *dX, double *dG, double *dsigma)
{
    double S[3];
    orc_stretching(transposed, J, G, S);
    double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    double csfs = C * sigma * sigma * sigma / zeta0;
    double e0 = E ? E[0] : 0.0, e1 = E ? E[1] : 0.0, e2 = E ? E[2] : 0.0;
    double SG = S[0] * G[0] + S[1] * G[1] + S[2] * G[2];
    double EG = e0 * G[0] + e1 * G[1] + e2 * G[2];
    double Z = 0.0;
    if (G2 > 0.0)
        Z = ((gcoef + f) / (1.0 + 3.0 * f) * SG - f / (1.0 + 3.0 * f) * csfs * EG) / G2;
    for (int k = 0; k < 3; k++) dX[k] = U[k] + Uinf[k];
    dG[0] = S[0] - 3.0 * Z * G[0] - csfs * e0;
    dG[1] = S[1] - 3.0 * Z * G[1] - csfs * e1;
    dG[2] = S[2] - 3.0 * Z * G[2] - csfs * e2;
    *dsigma = -sigma * Z;
}

/* Williamson (1980) low-storage 3-stage RK coefficients; the scheme is the
 * low-storage third-order RK of docs/src/theory/rvpm.md:364. */
static const double LSRK3_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double LSRK3_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

/* One step of the same scheme applied to y' = lambda y (complex lambda given as
 * re/im): the closed-form pin is 1 + z + z^2/2 + z^3/6 with z = lambda dt. */
void orc_lsrk3_linear(double lre, double lim, double dt, double *yre, double *yim)
{
    double qr = 0.0, qi = 0.0, yr = *yre, yi = *yim;
    for (int s = 0; s < 3; s++) {
        double fr = lre * yr - lim * yi, fi = lre * yi + lim * yr;
        qr = LSRK3_A[s] * qr + dt * fr;
        qi = LSRK3_A[s] * qi + dt * fi;
        yr += LSRK3_B[s] * qr;
        yi += LSRK3_B[s] * qi;
    }
    *yre = yr;
    *yim = yi;
}

/* Pedrizzetti relaxation (docs/src/theory/rvpm.md:363, Pedrizzetti 1992):
 *   Gamma <- (1 - rlxf) Gamma + rlxf |Gamma| omega / |omega|
 * with omega = curl U from J: (J[3,2]-J[2,3], J[1,3]-J[3,1], J[2,1]-J[1,2]).
 * kind 1 = pedrizzetti, kind 2 = corrected (rescaled to keep |Gamma|). */
void orc_relax(int kind, double rlxf, const double *J, double *G)
{
    if (kind == 0) return;
    double w1 = J[5] - J[7], w2 = J[6] - J[2], w3 = J[1] - J[3];
    double nw = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
    double nG = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
    if (nw == 0.0) return;
    double cosang = (G[0] * w1 + G[1] * w2 + G[2] * w3) / (nG * nw);
    G[0] = (1.0 - rlxf) * G[0] + rlxf * nG * w1 / nw;
    G[1] = (1.0 - rlxf) * G[1] + rlxf * nG * w2 / nw;
    G[2] = (1.0 - rlxf) * G[2] + rlxf * nG * w3 / nw;
    if (kind == 2) {
        double b2 = 1.0 - 2.0 * (1.0 - rlxf) * rlxf * (1.0 - cosang);
        double b = sqrt(b2);
        G[0] /= b; G[1] /= b; G[2] /= b;
    }
}

/* Particle field state for the integrators: AoS per particle. */

/* Core spreading's spatial adaptation (viscous_oracle.c): when a free particle's
 * core exceeds beta sgm0, every free particle's core is reset to sgm0 and its
 * strength re-fitted so that the free particles' vorticity at the free
 * particles is kept (static particles are left as they are and drop out of
 * both sides).  Returns 1 when a reset happened. */
/* The same with the FMM (P->fmm): every vorticity sum runs over the near
 * lists of one tree of all particles (statics included, with no strength)
 * built at the current positions and cores; conjugate gradients on the free
 * entries (static ones drop out of the products and keep their strength). */
static int core_spreading_reset_fmm(const orc_params *P, orc_field *F)
{
    const int64_t n = F->np;
    const uint8_t *mask = F->is_static;
    struct orc_near_plan *pl = orc_near_plan_build(P->ncrit, P->theta, P->sigcut, n, F->X,
                                                   F->sigma);
    double *w = malloc(sizeof(double) * 3 * n), *r = malloc(sizeof(double) * 3 * n),
           *p = malloc(sizeof(double) * 3 * n), *Ap = malloc(sizeof(double) * 3 * n);
    orc_near_zeta(pl, P->kernel, F->X, F->G, F->sigma, 0.0, mask, w);
    for (int64_t i = 0; i < n; i++)
        if (!(mask && mask[i])) F->sigma[i] = P->sgm0;
    orc_near_zeta(pl, P->kernel, F->X, F->G, NULL, P->sgm0, mask, Ap);
    double rr[3] = {0, 0, 0}, wn[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; i++) {
        const int off = mask && mask[i];
        for (int k = 0; k < 3; k++) {
            r[3 * i + k] = off ? 0.0 : w[3 * i + k] - Ap[3 * i + k];
            p[3 * i + k] = r[3 * i + k];
            if (!off) {
                rr[k] += r[3 * i + k] * r[3 * i + k];
                wn[k] += w[3 * i + k] * w[3 * i + k];
            }
        }
    }
    int it = 0, done[3];
    for (int k = 0; k < 3; k++) done[k] = !(sqrt(rr[k]) > P->tol * sqrt(wn[k]));
    while (it < P->itmax && !(done[0] && done[1] && done[2])) {
        orc_near_zeta(pl, P->kernel, F->X, p, NULL, P->sgm0, mask, Ap);
        double pAp[3] = {0, 0, 0}, rr1[3] = {0, 0, 0}, a[3];
        for (int64_t i = 0; i < n; i++)
            if (!(mask && mask[i]))
                for (int k = 0; k < 3; k++) pAp[k] += p[3 * i + k] * Ap[3 * i + k];
        for (int k = 0; k < 3; k++) a[k] = done[k] || pAp[k] == 0.0 ? 0.0 : rr[k] / pAp[k];
        for (int64_t i = 0; i < n; i++) {
            if (mask && mask[i]) continue;
            for (int k = 0; k < 3; k++) {
                F->G[3 * i + k] += a[k] * p[3 * i + k];
                r[3 * i + k] -= a[k] * Ap[3 * i + k];
                rr1[k] += r[3 * i + k] * r[3 * i + k];
            }
        }
        for (int k = 0; k < 3; k++) {
            const double b = done[k] || rr[k] == 0.0 ? 0.0 : rr1[k] / rr[k];
            if (!done[k])
                for (int64_t i = 0; i < n; i++)
                    if (!(mask && mask[i])) p[3 * i + k] = r[3 * i + k] + b * p[3 * i + k];
            rr[k] = rr1[k];
            done[k] = done[k] || !(sqrt(rr[k]) > P->tol * sqrt(wn[k]));
        }
        it++;
    }
    free(w); free(r); free(p); free(Ap);
    orc_near_plan_free(pl);
    return it;
}

int orc_core_spreading_reset(const orc_params *P, orc_field *F, int *iters)
{
    const int64_t n = F->np;
    int64_t nf = 0;
    double smax = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (F->is_static && F->is_static[i]) continue;
        nf++;
        smax = fmax(smax, F->sigma[i]);
    }
    if (iters) *iters = 0;
    if (nf == 0 || !(smax > P->beta * P->sgm0)) return 0;
    if (P->fmm) {
        const int it = core_spreading_reset_fmm(P, F);
        if (iters) *iters = it;
        return 1;
    }
    double *X = malloc(sizeof(double) * 3 * nf), *G = malloc(sizeof(double) * 3 * nf),
           *s = malloc(sizeof(double) * nf), *w = malloc(sizeof(double) * 3 * nf);
    int64_t *idx = malloc(sizeof(int64_t) * nf);
    int64_t q = 0;
    for (int64_t i = 0; i < n; i++) {
        if (F->is_static && F->is_static[i]) continue;
        idx[q] = i;
        for (int k = 0; k < 3; k++) {
            X[3 * q + k] = F->X[3 * i + k];
            G[3 * q + k] = F->G[3 * i + k];
        }
        s[q] = F->sigma[i];
        q++;
    }
    orc_zeta_sum(P->kernel, nf, X, nf, X, G, s, 0.0, w);
    const int it = orc_rbf_cg(P->kernel, nf, X, P->sgm0, w, G, P->itmax, P->tol, NULL);
    for (int64_t j = 0; j < nf; j++) {
        const int64_t i = idx[j];
        for (int k = 0; k < 3; k++) F->G[3 * i + k] = G[3 * j + k];
        F->sigma[i] = P->sgm0;
    }
    if (iters) *iters = it;
    free(X); free(G); free(s); free(w); free(idx);
    return 1;
}

static void field_uj(const orc_params *P, orc_field *F)
{
    if (P->fmm) {   /* overwrites U and J */
        orc_uj_fmm(P->kernel, P->p, P->ncrit, P->theta, P->sigcut, F->np, F->X, F->G, F->sigma,
                   F->U, F->J, NULL, 0, NULL);
        return;
    }
    memset(F->U, 0, sizeof(double) * 3 * F->np);
    memset(F->J, 0, sizeof(double) * 9 * F->np);
    orc_uj_direct(P->kernel, F->np, F->X, F->np, F->X, F->G, F->sigma, F->U, F->J);
}

/* Estr (and, with DJ, its core-size derivative DE) with the UJ's scheme. */
static void field_estr(const orc_params *P, orc_field *F, const double *DJ, double *DE)
{
    if (P->fmm) {
        orc_fmm_near(P->kernel, P->transposed, 2, P->ncrit, P->theta, P->sigcut, F->np, F->X,
                     F->G, F->sigma, F->J, DJ, F->E, DE);
        return;
    }
    memset(F->E, 0, sizeof(double) * 3 * F->np);
    if (DJ) {
        memset(DE, 0, sizeof(double) * 3 * F->np);
        orc_estr_dyn_direct(P->kernel, P->transposed, F->np, F->X, F->J, DJ, F->np, F->X, F->G,
                            F->sigma, F->J, DJ, F->E, DE);
    } else {
        orc_estr_direct(P->kernel, P->transposed, F->np, F->X, F->J, F->np, F->X, F->G, F->sigma,
                        F->J, F->E);
    }
}

/* Core-size derivatives D[U], D[J] with the UJ's scheme.  The FMM's far field
 * (singular kernel, multipoles) does not depend on sigma, so with the FMM
 * they are sums over its near-field lists only. */
static void field_dj(const orc_params *P, const orc_field *F, double *DU, double *DJ)
{
    if (P->fmm) {
        orc_fmm_near(P->kernel, P->transposed, 1, P->ncrit, P->theta, P->sigcut, F->np, F->X,
                     F->G, F->sigma, NULL, NULL, DU, DJ);
        return;
    }
    memset(DU, 0, sizeof(double) * 3 * F->np);
    memset(DJ, 0, sizeof(double) * 9 * F->np);
    orc_dj_direct(P->kernel, F->np, F->X, F->np, F->X, F->G, F->sigma, DU, DJ);
}

static int clipped(const orc_params *P, const orc_field *F, int64_t i, double c)
{
    const double *g = F->G + 3 * i, *e = F->E + 3 * i;
    return P->clip_backscatter && c * (g[0] * e[0] + g[1] * e[1] + g[2] * e[2]) < 0.0;
}

/* Numerator and denominator of the dynamic coefficient of one particle
 * (docs/src/theory/rvpm.md:264-290), C_d = <Gamma.L> / <Gamma.m> with
 *   m = sigma^3/zeta0 dE/dsigma,
 *   L = (3/sigma) (Gamma.grad)(u - ubar) + (Gamma.grad) dubar/dsigma.
 * The sigma-derivatives are exact (orc_kernel_dsig): dubar/dsigma = D[ubar]/sigma
 * and dE/dsigma = D[E]/sigma with D = sigma d/dsigma over every core at once.
 * The unfiltered velocity u is not resolved; the procedure's test filter
 * alpha sigma (alpha < 1, a finer filter) stands in for it, as the third level
 * of a pseudo-three-level procedure, to first order in 1 - alpha:
 *   u - ubar ~ ubar(alpha sigma) - ubar(sigma) ~ (alpha - 1) D[ubar],
 * so L = (3 alpha - 2)/sigma (Gamma.grad) D[ubar] and
 *   num = (3 alpha - 2)/sigma  Gamma^T D[J] Gamma   (the same for J or J^T),
 *   den = sigma^2/zeta0  Gamma . D[E]. */
void orc_dyn_terms(double alpha, double zeta0, double sigma, const double *G, const double *DJ,
                   const double *DE, double *num, double *den)
{
    double q = 0.0;
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) q += G[k] * DJ[k + 3 * l] * G[l];
    *num = (3.0 * alpha - 2.0) / sigma * q;
    *den = sigma * sigma / zeta0 * (G[0] * DE[0] + G[1] * DE[1] + G[2] * DE[2]);
}

/* One stage's UJ and SFS coefficient.
 * Constant model: E = Estr, C = Cs, then backscatter clipping C = 0 where
 * C Gamma.E < 0 (docs/src/theory/rvpm.md:296).
 * Dynamic model: num and den of orc_dyn_terms, each averaged along the
 * particle's trajectory as a <- w a + (1 - w) a_old (the first sample is
 * taken as is), C = <num>/<den> bounded to [minC, maxC], then backscatter
 * clipping.  FLOWVPM's own implementation is not in the reference tree; this
 * restates the documented procedure with both terms of L. */
static void field_stage(const orc_params *P, orc_field *F)
{
    const int64_t n = F->np;
    field_uj(P, F);
    if (!P->sfs) return;
    if (P->sfs == 1) {
        field_estr(P, F, NULL, NULL);
        for (int64_t i = 0; i < n; i++) {
            double c = P->Cs;
            if (clipped(P, F, i, c)) c = 0.0;
            F->C[3 * i] = c;
        }
        return;
    }
    const size_t m = n > 0 ? (size_t)n : 1;
    double *DU = malloc(sizeof(double) * 3 * m), *DJ = malloc(sizeof(double) * 9 * m),
           *DE = malloc(sizeof(double) * 3 * m);
    field_dj(P, F, DU, DJ);
    field_estr(P, F, DJ, DE);
    const double zeta0 = orc_zeta0(P->kernel);
    for (int64_t i = 0; i < n; i++) {
        if (F->is_static && F->is_static[i]) continue;
        double num, den;
        orc_dyn_terms(P->alpha, zeta0, F->sigma[i], F->G + 3 * i, DJ + 9 * i, DE + 3 * i, &num,
                      &den);
        double *C = F->C + 3 * i;
        double na = num, da = den;
        if (C[1] != 0.0 || C[2] != 0.0) {
            na = P->sfs_rlxf * num + (1.0 - P->sfs_rlxf) * C[1];
            da = P->sfs_rlxf * den + (1.0 - P->sfs_rlxf) * C[2];
        }
        C[1] = na;
        C[2] = da;
        double c = da != 0.0 ? na / da : 0.0;
        c = fmin(fmax(c, P->minC), P->maxC);
        if (clipped(P, F, i, c)) c = 0.0;
        C[0] = c;
    }
    free(DU); free(DJ); free(DE);
}

static void field_relax(const orc_params *P, orc_field *F)
{
    field_uj(P, F);
    for (int64_t i = 0; i < F->np; i++) {
        if (F->is_static && F->is_static[i]) continue;
        orc_relax(P->relax_kind, P->rlxf, F->J + 9 * i, F->G + 3 * i);
    }
}

/* rungekutta3: three low-storage stages of (UJ, SFS, update of non-static
 * particles), then optional relaxation after a fresh UJ.  Stage structure per
 * vpm.nextstep (src/FLOWUnsteady_simulation.jl:358); relax cadence per
 * src/FLOWUnsteady_simulation.jl:346-348. */
/* Viscous part of one stage's right-hand side: PSE adds to dGamma/dt
 * (computed on the stage's state), core spreading adds nu / sigma to
 * dsigma/dt.  V (n x 3) holds the PSE term. */
static double *viscous_stage(const orc_params *P, const orc_field *F)
{
    if (P->viscous != 2) return NULL;
    double *V = malloc(sizeof(double) * 3 * (F->np > 0 ? F->np : 1));
    if (P->fmm) {   /* over the near lists of the stage's FMM tree */
        struct orc_near_plan *pl = orc_near_plan_build(P->ncrit, P->theta, P->sigcut, F->np,
                                                       F->X, F->sigma);
        orc_near_pse(pl, P->kernel, P->nu, F->X, F->G, F->sigma, F->vol, V);
        orc_near_plan_free(pl);
        return V;
    }
    orc_pse(P->kernel, P->nu, F->np, F->X, F->G, F->sigma, F->vol, F->np, NULL, V);
    return V;
}

static void add_viscous(const orc_params *P, const double *V, int64_t i, double sigma,
                        double *dG, double *ds)
{
    if (P->viscous == 2)
        for (int k = 0; k < 3; k++) dG[k] += V[3 * i + k];
    else if (P->viscous == 1)
        *ds += P->nu / sigma;
}

void orc_rk3_step(const orc_params *P, orc_field *F, double dt, int relax)
{
    double zeta0 = orc_zeta0(P->kernel);
    memset(F->M, 0, sizeof(double) * 7 * F->np);
    for (int s = 0; s < 3; s++) {
        double a = LSRK3_A[s], b = LSRK3_B[s];
        field_stage(P, F);
        double *V = viscous_stage(P, F);
        for (int64_t i = 0; i < F->np; i++) {
            if (F->is_static && F->is_static[i]) continue;
            double dX[3], dG[3], ds;
            orc_rhs(P->transposed, P->f, P->g, zeta0, P->Uinf, F->U + 3 * i, F->J + 9 * i,
                    F->G + 3 * i, F->sigma[i], P->sfs ? F->C[3 * i] : 0.0,
                    P->sfs ? F->E + 3 * i : NULL, dX, dG, &ds);
            add_viscous(P, V, i, F->sigma[i], dG, &ds);
            double *M = F->M + 7 * i;
            for (int k = 0; k < 3; k++) {
                M[k] = a * M[k] + dt * dX[k];
                F->X[3 * i + k] += b * M[k];
            }
            for (int k = 0; k < 3; k++) {
                M[3 + k] = a * M[3 + k] + dt * dG[k];
                F->G[3 * i + k] += b * M[3 + k];
            }
            M[6] = a * M[6] + dt * ds;
            F->sigma[i] += b * M[6];
        }
        free(V);
    }
    if (relax && P->relax_kind) field_relax(P, F);
    if (P->viscous == 1) orc_core_spreading_reset(P, F, NULL);
}

/* euler: one UJ (+SFS), forward-Euler update, then relaxation with that same J
 * when requested. */
void orc_euler_step(const orc_params *P, orc_field *F, double dt, int relax)
{
    double zeta0 = orc_zeta0(P->kernel);
    field_stage(P, F);
    double *V = viscous_stage(P, F);
    for (int64_t i = 0; i < F->np; i++) {
        if (F->is_static && F->is_static[i]) continue;
        double dX[3], dG[3], ds;
        orc_rhs(P->transposed, P->f, P->g, zeta0, P->Uinf, F->U + 3 * i, F->J + 9 * i,
                F->G + 3 * i, F->sigma[i], P->sfs ? F->C[3 * i] : 0.0, P->sfs ? F->E + 3 * i : NULL,
                dX, dG, &ds);
        add_viscous(P, V, i, F->sigma[i], dG, &ds);
        for (int k = 0; k < 3; k++) {
            F->X[3 * i + k] += dt * dX[k];
            F->G[3 * i + k] += dt * dG[k];
        }
        F->sigma[i] += dt * ds;
        if (relax && P->relax_kind)
            orc_relax(P->relax_kind, P->rlxf, F->J + 9 * i, F->G + 3 * i);
    }
    free(V);
    if (P->viscous == 1) orc_core_spreading_reset(P, F, NULL);
}

/* Enstrophy 0.5 sum_p Gamma_p . omega(x_p), omega from J
 * (src/FLOWUnsteady_monitors.jl:573-574). */
double orc_enstrophy(int64_t np, const double *G, const double *J)
{
    double s = 0.0;
    for (int64_t i = 0; i < np; i++) {
        const double *j = J + 9 * i, *g = G + 3 * i;
        s += g[0] * (j[5] - j[7]) + g[1] * (j[6] - j[2]) + g[2] * (j[1] - j[3]);
    }
    return 0.5 * s;
}

/* Kinetic energy in Saffman's impulse form, E = sum_p U(x_p) . (x_p x Gamma_p).
 * The reference defines no energy monitor; this definition is this oracle's. */
double orc_energy(int64_t np, const double *X, const double *G, const double *U)
{
    double s = 0.0;
    for (int64_t i = 0; i < np; i++) {
        const double *x = X + 3 * i, *g = G + 3 * i, *u = U + 3 * i;
        double c0 = x[1] * g[2] - x[2] * g[1], c1 = x[2] * g[0] - x[0] * g[2],
               c2 = x[0] * g[1] - x[1] * g[0];
        s += u[0] * c0 + u[1] * c1 + u[2] * c2;
    }
    return s;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
