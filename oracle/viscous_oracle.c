/*
 * viscous_oracle.c -- CPU FP64 restatement of the viscous diffusion schemes
 * FLOWUnsteady selects with vpm_viscous (src/FLOWUnsteady_simulation.jl:42,133,
 * 229-234): vpm.ParticleStrengthExchange(nu) and vpm.CoreSpreading(nu, sgm0,
 * zeta; beta, itmax, tol) with radial-basis-function spatial adaptation solved
 * by conjugate gradients (vpm.rbf_conjugategradient; docs/src/api/flowvpm-viscous.md).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; the product never links it).
 *
 * FLOWVPM's implementation of these schemes is not in the reference tree, so
 * this restates the published methods the reference cites for the viscous term
 * nu lap(omega) of docs/src/theory/rvpm.md:182-193,367-368:
 *   PSE (Degond & Mas-Gallic 1989): with eta(rho) = -zeta'(rho)/rho, whose
 *     second moment int x_1^2 eta = int zeta = 1,
 *       dGamma_i/dt = (2 nu / s^2) sum_j (vol_i Gamma_j - vol_j Gamma_i) eta_s(x_i - x_j),
 *     eta_s(r) = s^-3 eta(r/s), s^2 = (sigma_i^2 + sigma_j^2)/2 (symmetric, so the
 *     sum conserves sum_i Gamma_i exactly).
 *   Core spreading (Barba, Leonard & Allen 2005, rvpm.md:367,390): Gaussian
 *     cores solve the heat equation when sigma^2 grows by 2 nu per unit time,
 *     d sigma/dt = nu / sigma; spatial adaptation resets every core to sgm0 and
 *     finds new strengths Gamma' by radial basis function interpolation of the
 *     vorticity at the particles,
 *       sum_j Gamma'_j zeta_sgm0(x_i - x_j) = omega(x_i) = sum_j Gamma_j zeta_sigma_j(x_i - x_j),
 *     solved by conjugate gradients (the matrix is symmetric positive definite).
 * Parity with FLOWVPM's own schemes is unpinned; the pins here are the
 * methods' defining properties (tests/test_viscous_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_kernel(int kernel, double rho, double *zeta, double *g, double *dg);

enum { V_SINGULAR = 0, V_GAUSSIAN = 1, V_GAUSSIANERF = 2, V_WINCKELMANS = 3 };

/* eta(rho) = -zeta'(rho)/rho of each smooth kernel (NAN for the singular one). */
double orc_eta(int kernel, double rho)
{
    double z;
    orc_kernel(kernel, rho, &z, NULL, NULL);
    switch (kernel) {
    case V_GAUSSIAN: return 3.0 * rho * z;            /* zeta = 3/(4 pi) e^{-rho^3} */
    case V_GAUSSIANERF: return z;                     /* zeta ~ e^{-rho^2/2} */
    case V_WINCKELMANS: return 7.0 * z / (rho * rho + 1.0);   /* zeta ~ (rho^2+1)^{-7/2} */
    default: return NAN;
    }
}

/* PSE right-hand side at the nt particles listed in tidx (every particle, in
 * order, when tidx is NULL): dG (nt x 3) is overwritten. */
void orc_pse(int kernel, double nu, int64_t n, const double *X, const double *G,
             const double *sig, const double *vol, int64_t nt, const int64_t *tidx, double *dG)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < nt; t++) {
        const int64_t i = tidx ? tidx[t] : t;
        double acc[3] = {0, 0, 0};
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            const double d0 = X[3 * i] - X[3 * j], d1 = X[3 * i + 1] - X[3 * j + 1],
                         d2 = X[3 * i + 2] - X[3 * j + 2];
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double s2 = 0.5 * (sig[i] * sig[i] + sig[j] * sig[j]);
            const double s = sqrt(s2);
            const double w = 2.0 * nu / s2 * orc_eta(kernel, r / s) / (s2 * s);
            for (int k = 0; k < 3; k++) acc[k] += w * (vol[i] * G[3 * j + k] - vol[j] * G[3 * i + k]);
        }
        for (int k = 0; k < 3; k++) dG[3 * t + k] = acc[k];
    }
}

/* Vorticity of the regularised particles at nt points: w (nt x 3) overwritten,
 * w_i = sum_j Gamma_j zeta(r_ij / s_j) / s_j^3 with s_j = ssig[j], or s0 for
 * every source when s0 > 0.  Coincident points are included (zeta(0)). */
void orc_zeta_sum(int kernel, int64_t nt, const double *tX, int64_t ns, const double *sX,
                  const double *sG, const double *ssig, double s0, double *w)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; i++) {
        double acc[3] = {0, 0, 0};
        for (int64_t j = 0; j < ns; j++) {
            const double d0 = tX[3 * i] - sX[3 * j], d1 = tX[3 * i + 1] - sX[3 * j + 1],
                         d2 = tX[3 * i + 2] - sX[3 * j + 2];
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double s = s0 > 0.0 ? s0 : ssig[j];
            double z;
            orc_kernel(kernel, r / s, &z, NULL, NULL);
            z /= s * s * s;
            for (int k = 0; k < 3; k++) acc[k] += z * sG[3 * j + k];
        }
        for (int k = 0; k < 3; k++) w[3 * i + k] = acc[k];
    }
}

/* Conjugate gradients on A G = w, A_ij = zeta_s0(x_i - x_j), one independent
 * iteration per component, from the initial guess in G (n x 3, overwritten),
 * until ||r_k|| <= tol ||w_k|| for every component or itmax iterations.
 * Returns the iterations taken; res[k] = ||r_k|| / ||w_k|| at the end. */
int orc_rbf_cg(int kernel, int64_t n, const double *X, double s0, const double *w, double *G,
               int itmax, double tol, double *res)
{
    double *r = malloc(sizeof(double) * 3 * n), *p = malloc(sizeof(double) * 3 * n),
           *Ap = malloc(sizeof(double) * 3 * n);
    double rr[3], wn[3];
    orc_zeta_sum(kernel, n, X, n, X, G, NULL, s0, Ap);
    for (int k = 0; k < 3; k++) rr[k] = wn[k] = 0.0;
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            r[3 * i + k] = w[3 * i + k] - Ap[3 * i + k];
            p[3 * i + k] = r[3 * i + k];
            rr[k] += r[3 * i + k] * r[3 * i + k];
            wn[k] += w[3 * i + k] * w[3 * i + k];
        }
    int it = 0;
    int done[3];
    for (int k = 0; k < 3; k++) done[k] = !(sqrt(rr[k]) > tol * sqrt(wn[k]));
    while (it < itmax && !(done[0] && done[1] && done[2])) {
        orc_zeta_sum(kernel, n, X, n, X, p, NULL, s0, Ap);
        double pAp[3] = {0, 0, 0};
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) pAp[k] += p[3 * i + k] * Ap[3 * i + k];
        double rr1[3] = {0, 0, 0}, a[3];
        for (int k = 0; k < 3; k++) a[k] = done[k] || pAp[k] == 0.0 ? 0.0 : rr[k] / pAp[k];
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                G[3 * i + k] += a[k] * p[3 * i + k];
                r[3 * i + k] -= a[k] * Ap[3 * i + k];
                rr1[k] += r[3 * i + k] * r[3 * i + k];
            }
        for (int k = 0; k < 3; k++) {
            const double b = done[k] || rr[k] == 0.0 ? 0.0 : rr1[k] / rr[k];
            if (!done[k])
                for (int64_t i = 0; i < n; i++) p[3 * i + k] = r[3 * i + k] + b * p[3 * i + k];
            rr[k] = rr1[k];
            done[k] = done[k] || !(sqrt(rr[k]) > tol * sqrt(wn[k]));
        }
        it++;
    }
    if (res)
        for (int k = 0; k < 3; k++) res[k] = wn[k] > 0.0 ? sqrt(rr[k] / wn[k]) : 0.0;
    free(r);
    free(p);
    free(Ap);
    return it;
}
