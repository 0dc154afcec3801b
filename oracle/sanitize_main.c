/*
 * sanitize_main.c -- drives every oracle entry point on small inputs so that
 * an AddressSanitizer / UndefinedBehaviorSanitizer build of the oracle
 * (`make -C oracle sanitize`) can check it.  TEST INFRASTRUCTURE ONLY.
 * Exits 0 when every call returned and every result is finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

static double urand(unsigned *s)
{
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0;
}

static int finite_all(const double *a, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        if (!isfinite(a[i])) return 0;
    return 1;
}

#define CHECK(a, n, what)                                 \
    do {                                                  \
        if (!finite_all((a), (n))) {                      \
            fprintf(stderr, "non-finite: %s\n", (what));  \
            return 1;                                     \
        }                                                 \
    } while (0)

int main(void)
{
    enum { N = 300, NB = 7 };
    unsigned seed = 12345u;
    double *X = malloc(sizeof(double) * 3 * N), *G = malloc(sizeof(double) * 3 * N);
    double *sg = malloc(sizeof(double) * N), *vol = malloc(sizeof(double) * N);
    double *U = calloc(3 * N, sizeof(double)), *J = calloc(9 * N, sizeof(double));
    double *DU = calloc(3 * N, sizeof(double)), *DJ = calloc(9 * N, sizeof(double));
    double *E = calloc(3 * N, sizeof(double)), *DE = calloc(3 * N, sizeof(double));
    double *M = calloc(7 * N, sizeof(double)), *C = calloc(3 * N, sizeof(double));
    double *w = calloc(3 * N, sizeof(double)), *V = calloc(3 * N, sizeof(double));
    uint8_t *st = calloc(N, 1);
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < 3; k++) {
            X[3 * i + k] = 2.0 * urand(&seed) - 1.0;
            G[3 * i + k] = 1e-2 * (urand(&seed) - 0.5);
        }
        sg[i] = 0.1 + 0.1 * urand(&seed);
        vol[i] = sg[i] * sg[i] * sg[i];
        st[i] = (uint8_t)(i % 17 == 0);
    }
    for (int kern = 0; kern < 4; kern++) {
        double z, g, dg, h, dh, zD;
        orc_kernel(kern, 0.7, &z, &g, &dg);
        orc_kernel_dsig(kern, 0.7, &h, &dh, &zD);
        (void)orc_dzeta(kern, 0.7);
        (void)orc_zeta0(kern);
        memset(U, 0, sizeof(double) * 3 * N);
        memset(J, 0, sizeof(double) * 9 * N);
        orc_uj_direct(kern, N, X, N, X, G, sg, U, J);
        CHECK(U, 3 * N, "uj_direct U");
        CHECK(J, 9 * N, "uj_direct J");
        orc_uj_fmm(kern, 4, 20, 0.5, kern == 0 ? 0.0 : 3.0, N, X, G, sg, U, J, NULL, 0, NULL);
        CHECK(J, 9 * N, "uj_fmm J");
        if (kern == 0) continue;
        memset(DU, 0, sizeof(double) * 3 * N);
        memset(DJ, 0, sizeof(double) * 9 * N);
        orc_dj_direct(kern, N, X, N, X, G, sg, DU, DJ);
        CHECK(DJ, 9 * N, "dj_direct");
        memset(E, 0, sizeof(double) * 3 * N);
        memset(DE, 0, sizeof(double) * 3 * N);
        orc_estr_direct(kern, 1, N, X, J, N, X, G, sg, J, E);
        orc_estr_dyn_direct(kern, 0, N, X, J, DJ, N, X, G, sg, J, DJ, E, DE);
        CHECK(DE, 3 * N, "estr_dyn_direct");
        if (orc_fmm_near(kern, 1, 1, 20, 0.5, 0.0, N, X, G, sg, NULL, NULL, DU, DJ) ||
            orc_fmm_near(kern, 1, 2, 20, 0.5, 0.0, N, X, G, sg, J, DJ, E, DE) ||
            orc_fmm_near(kern, 0, 2, 20, 0.5, 0.0, N, X, G, sg, J, NULL, E, NULL))
            return 1;
        CHECK(DE, 3 * N, "fmm_near DE");
        double num, den;
        orc_dyn_terms(0.999, orc_zeta0(kern), sg[0], G, DJ, DE, &num, &den);
    }
    double out[6 * NB];
    orc_ffv_direct(2, NB, X, G, sg, N - NB, X + 3 * NB, G + 3 * NB, out);
    CHECK(out, 6 * NB, "ffv_direct");
    orc_pse(2, 1e-3, N, X, G, sg, vol, N, NULL, V);
    orc_zeta_sum(2, N, X, N, X, G, sg, 0.0, w);
    memcpy(V, G, sizeof(double) * 3 * N);
    (void)orc_rbf_cg(2, N, X, 0.12, w, V, 5, 1e-6, NULL);
    CHECK(V, 3 * N, "rbf_cg");
    double yr = 1.0, yi = 0.0, S[3], dX[3], dG[3], ds, Uinf[3] = {0.1, 0.0, 0.0};
    orc_lsrk3_linear(-1.0, 2.0, 0.1, &yr, &yi);
    orc_stretching(1, J, G, S);
    orc_rhs(1, 0.0, 0.2, orc_zeta0(2), Uinf, U, J, G, sg[0], 0.1, E, dX, dG, &ds);
    orc_relax(2, 0.3, J, V);
    (void)orc_enstrophy(N, G, J);
    (void)orc_energy(N, X, G, U);
    (void)orc_num_threads();
    int mi[3 * 35];
    double b[35], r[3] = {0.3, -0.2, 0.5};
    orc_fmm_deriv(4, r, b, mi);

    /* integrators: direct and FMM, dynamic SFS, static particles, PSE, core spreading */
    orc_field F = {N, X, G, sg, U, J, E, M, C, st, vol};
    for (int fmm = 0; fmm < 2; fmm++)
        for (int visc = 0; visc < 3; visc++) {
            orc_params P;
            memset(&P, 0, sizeof(P));
            P.kernel = 2; P.transposed = 1; P.g = 0.2;
            P.sfs = 2; P.alpha = 0.999; P.sfs_rlxf = 0.005; P.maxC = 1.0; P.clip_backscatter = 1;
            P.relax_kind = 1; P.rlxf = 0.3;
            P.viscous = visc; P.nu = 1e-3; P.sgm0 = 0.1; P.beta = 1.0; P.itmax = 5; P.tol = 1e-6;
            P.fmm = fmm; P.p = 4; P.ncrit = 20; P.theta = 0.5;
            orc_rk3_step(&P, &F, 1e-3, 1);
            orc_euler_step(&P, &F, 1e-3, 1);
            CHECK(X, 3 * N, "step X");
            CHECK(G, 3 * N, "step Gamma");
            CHECK(C, 3 * N, "step C");
        }
    int it = 0;
    orc_params P;
    memset(&P, 0, sizeof(P));
    P.kernel = 2; P.sgm0 = 0.05; P.beta = 0.5; P.itmax = 3; P.tol = 1e-6;
    (void)orc_core_spreading_reset(&P, &F, &it);
    free(X); free(G); free(sg); free(vol); free(U); free(J); free(DU); free(DJ);
    free(E); free(DE); free(M); free(C); free(w); free(V); free(st);
    printf("oracle sanitize run: ok\n");
    return 0;
}
