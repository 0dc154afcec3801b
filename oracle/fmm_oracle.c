/*
 * fmm_oracle.c -- CPU FP64 fast multipole method for U and J (UJ_fmm).
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the host-core FMM baseline; the
 * product never links it).
 *
 * Restates what the reference's UJ_fmm computes (src/FLOWUnsteady_simulation.jl:38,43,129;
 * docs/src/theory/rvpm.md:369-370): the regularised Biot-Savart U and J of
 * all particles with an O(N) tree code controlled by p, ncrit and theta.  The
 * FMM backend itself (a modified ExaFMM) is not in the reference tree, so the
 * algorithm here is this build's own:
 *
 *  - vector potential psi = (1/4 pi) sum_j Gamma_j / |x - x_j| far from the
 *    sources; u = curl psi, J = grad u, three Laplace potentials with charges
 *    Gamma_j,k;
 *  - Cartesian Taylor expansions truncated at total degree p: raw multipole
 *    moments m_a = sum q (x_j - c)^a, locals phi(x) = sum_b l_b (x - a)^b,
 *    derivatives of 1/r through the recurrence
 *      R^2 n b_k = -(2n-1) sum_i r_i b_{k-e_i} - (n-1) sum_i b_{k-2e_i},
 *    b_k = (d^k (1/r)) / k!, and M2L
 *      l_b += (-1)^{|a|} C(a+b, a) m_a b_{a+b}(a_center - c_center),  |a|+|b| <= p;
 *  - adaptive octree over 63-bit Morton keys (21 bits per axis) of the
 *    bounding cube, leaves of at most ncrit particles, geometric cell radius;
 *  - dual tree traversal: a pair (target A, source B) is accepted for M2L when
 *    (R_A + R_B) < theta |c_A - c_B| and the gap |c_A - c_B| - R_A - R_B is at
 *    least sigcut * max sigma in B (so the singular far field is exact to the
 *    kernel's tail); otherwise the larger (or non-leaf) cell is split, and
 *    leaf-leaf pairs become direct regularised P2P (orc_pair_uj).
 *
 * Pins (tests/test_fmm_oracle.py): the error against orc_uj_direct falls with
 * p; theta -> 0 reproduces the direct sum; a single far source is reproduced
 * to the closed-form Taylor remainder.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_CONST4 0.07957747154594767
#define MAXP 16
#define KEYBITS 21

void orc_pair_uj(int kernel, const double *xt, const double *xs, const double *gs, double sgm,
                 double *U, double *J);

/* ---- multi-indices -------------------------------------------------------------- */
typedef struct {
    int P, K;                     /* max degree, count of |a| <= P */
    int (*mi)[3];
    int *lut;                     /* lut[(a*(P+1)+b)*(P+1)+c] */
} mindex;

static void mi_build(mindex *M, int P)
{
    M->P = P;
    M->K = (P + 1) * (P + 2) * (P + 3) / 6;
    M->mi = malloc(sizeof(int[3]) * M->K);
    M->lut = malloc(sizeof(int) * (P + 1) * (P + 1) * (P + 1));
    for (int i = 0; i < (P + 1) * (P + 1) * (P + 1); i++) M->lut[i] = -1;
    int k = 0;
    for (int n = 0; n <= P; n++)
        for (int a = n; a >= 0; a--)
            for (int b = n - a; b >= 0; b--) {
                int c = n - a - b;
                M->mi[k][0] = a; M->mi[k][1] = b; M->mi[k][2] = c;
                M->lut[(a * (P + 1) + b) * (P + 1) + c] = k;
                k++;
            }
}

static void mi_free(mindex *M) { free(M->mi); free(M->lut); }

static inline int mi_idx(const mindex *M, int a, int b, int c)
{
    if (a < 0 || b < 0 || c < 0 || a + b + c > M->P) return -1;
    return M->lut[(a * (M->P + 1) + b) * (M->P + 1) + c];
}

/* b_k(r) = (1/k!) d^k (1/|r|), |k| <= P, by the recurrence in the header. */
static void deriv_b(const mindex *M, const double r[3], double *b)
{
    double R2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    b[0] = 1.0 / sqrt(R2);
    for (int k = 1; k < M->K; k++) {
        const int *a = M->mi[k];
        int n = a[0] + a[1] + a[2];
        double s1 = 0.0, s2 = 0.0;
        for (int i = 0; i < 3; i++) {
            int d1[3] = {a[0], a[1], a[2]};
            d1[i] -= 1;
            int j1 = mi_idx(M, d1[0], d1[1], d1[2]);
            if (j1 >= 0) s1 += r[i] * b[j1];
            d1[i] -= 1;
            int j2 = mi_idx(M, d1[0], d1[1], d1[2]);
            if (j2 >= 0) s2 += b[j2];
        }
        b[k] = -((2.0 * n - 1.0) * s1 + (n - 1.0) * s2) / (n * R2);
    }
}

/* t^a for |a| <= P */
static void monomials(const mindex *M, const double t[3], double *out)
{
    double pw[3][MAXP + 3];
    for (int i = 0; i < 3; i++) {
        pw[i][0] = 1.0;
        for (int e = 1; e <= M->P; e++) pw[i][e] = pw[i][e - 1] * t[i];
    }
    for (int k = 0; k < M->K; k++) out[k] = pw[0][M->mi[k][0]] * pw[1][M->mi[k][1]] * pw[2][M->mi[k][2]];
}

static double binom(int n, int k)
{
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

/* ---- tree ---------------------------------------------------------------------- */
typedef struct {
    int64_t begin, end;          /* particle range in sorted order */
    int64_t child0;              /* first child cell (children contiguous), -1 if leaf */
    int nchild, level;
    double c[3], R, smax;
} orc_cell;

typedef struct {
    int64_t n, ncell, cap;
    orc_cell *cells;
    uint64_t *key;               /* sorted keys */
    int64_t *perm;               /* sorted slot -> original index */
    double lo[3], ext;
} orc_tree;

typedef struct { uint64_t k; int64_t i; } keyidx;

static int cmp_keyidx(const void *a, const void *b)
{
    const keyidx *x = a, *y = b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}

static uint64_t spread21(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

/* Morton key of a position: 21 bits per axis of the bounding cube. */
uint64_t orc_morton63(const double x[3], const double lo[3], double ext)
{
    const double scale = 2097152.0 / ext;
    uint64_t q[3];
    for (int k = 0; k < 3; k++) {
        double v = (x[k] - lo[k]) * scale;
        if (v < 0.0) v = 0.0;
        if (v > 2097151.0) v = 2097151.0;
        q[k] = (uint64_t)v;
    }
    return spread21(q[0]) | (spread21(q[1]) << 1) | (spread21(q[2]) << 2);
}

static int64_t lower_bound_key(const uint64_t *key, int64_t b, int64_t e, uint64_t v)
{
    while (b < e) {
        int64_t m = b + (e - b) / 2;
        if (key[m] < v) b = m + 1; else e = m;
    }
    return b;
}

static void cell_geometry(orc_tree *T, orc_cell *C, uint64_t prefix)
{
    /* de-interleave prefix (3*level bits) into cell integer coordinates */
    uint64_t ix = 0, iy = 0, iz = 0;
    for (int l = 0; l < C->level; l++) {
        uint64_t o = (prefix >> (3 * (C->level - 1 - l))) & 7u;
        ix = (ix << 1) | (o & 1u);
        iy = (iy << 1) | ((o >> 1) & 1u);
        iz = (iz << 1) | ((o >> 2) & 1u);
    }
    double h = T->ext / (double)((uint64_t)1 << C->level);
    C->c[0] = T->lo[0] + ((double)ix + 0.5) * h;
    C->c[1] = T->lo[1] + ((double)iy + 0.5) * h;
    C->c[2] = T->lo[2] + ((double)iz + 0.5) * h;
    C->R = 0.8660254037844386 * h;
}

/* Builds the tree in breadth-first order (children contiguous, levels
 * contiguous).  X is AoS (n,3). */
static void tree_build(orc_tree *T, int64_t n, const double *X, const double *sig, int ncrit)
{
    T->n = n;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], X[3 * i + k]);
            hi[k] = fmax(hi[k], X[3 * i + k]);
        }
    double ext = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
    if (!(ext > 0.0)) ext = 1.0;
    ext *= 1.0000001;
    for (int k = 0; k < 3; k++) T->lo[k] = lo[k];
    T->ext = ext;
    keyidx *ki = malloc(sizeof(keyidx) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        ki[i].k = orc_morton63(X + 3 * i, lo, ext);
        ki[i].i = i;
    }
    qsort(ki, n, sizeof(keyidx), cmp_keyidx);
    T->key = malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    T->perm = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) { T->key[i] = ki[i].k; T->perm[i] = ki[i].i; }
    free(ki);

    T->cap = 1024;
    T->cells = malloc(sizeof(orc_cell) * T->cap);
    uint64_t *prefix = malloc(sizeof(uint64_t) * T->cap);
    T->ncell = 1;
    orc_cell *root = &T->cells[0];
    root->begin = 0; root->end = n; root->level = 0; root->child0 = -1; root->nchild = 0;
    prefix[0] = 0;
    cell_geometry(T, root, 0);
    for (int64_t ci = 0; ci < T->ncell; ci++) {
        orc_cell C = T->cells[ci];
        if (C.end - C.begin <= ncrit || C.level >= KEYBITS) continue;
        int shift = 3 * (KEYBITS - C.level - 1);
        int64_t first = T->ncell;
        int nch = 0;
        for (uint64_t o = 0; o < 8; o++) {
            uint64_t cp = (prefix[ci] << 3) | o;
            int64_t b = lower_bound_key(T->key, C.begin, C.end, cp << shift);
            int64_t e = lower_bound_key(T->key, C.begin, C.end, (cp + 1) << shift);
            if (e == b) continue;
            if (T->ncell == T->cap) {
                T->cap *= 2;
                T->cells = realloc(T->cells, sizeof(orc_cell) * T->cap);
                prefix = realloc(prefix, sizeof(uint64_t) * T->cap);
            }
            orc_cell *D = &T->cells[T->ncell];
            D->begin = b; D->end = e; D->level = C.level + 1; D->child0 = -1; D->nchild = 0;
            prefix[T->ncell] = cp;
            cell_geometry(T, D, cp);
            T->ncell++;
            nch++;
        }
        T->cells[ci].child0 = first;
        T->cells[ci].nchild = nch;
    }
    free(prefix);
    /* max sigma per cell, bottom-up (children follow parents in BFS order) */
    for (int64_t ci = T->ncell - 1; ci >= 0; ci--) {
        orc_cell *C = &T->cells[ci];
        double s = 0.0;
        if (C->nchild == 0) {
            for (int64_t j = C->begin; j < C->end; j++) s = fmax(s, sig[T->perm[j]]);
        } else {
            for (int c = 0; c < C->nchild; c++) s = fmax(s, T->cells[C->child0 + c].smax);
        }
        C->smax = s;
    }
}

static void tree_free(orc_tree *T)
{
    free(T->cells); free(T->key); free(T->perm);
}

/* ---- traversal ------------------------------------------------------------------------ */
typedef struct { int64_t *a; int64_t n, cap; } pairlist;   /* pairs (t, s) flattened */

static void pl_push(pairlist *L, int64_t t, int64_t s)
{
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 1024;
        L->a = realloc(L->a, sizeof(int64_t) * 2 * L->cap);
    }
    L->a[2 * L->n] = t;
    L->a[2 * L->n + 1] = s;
    L->n++;
}

static int cmp_pair(const void *x, const void *y)
{
    const int64_t *a = x, *b = y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    return a[1] < b[1] ? -1 : (a[1] > b[1]);
}

/* Multipole acceptance: theta criterion plus a gap of sigcut * sigma_max(B). */
int orc_fmm_mac(const double cA[3], double RA, const double cB[3], double RB, double smaxB,
                double theta, double sigcut)
{
    double dx = cA[0] - cB[0], dy = cA[1] - cB[1], dz = cA[2] - cB[2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    return (RA + RB) < theta * d && (d - RA - RB) >= sigcut * smaxB;
}

static void traverse(const orc_tree *T, double theta, double sigcut, pairlist *m2l, pairlist *p2p)
{
    int64_t cap = 1 << 16, top = 0;
    int64_t *st = malloc(sizeof(int64_t) * 2 * cap);
    st[0] = 0; st[1] = 0; top = 1;
    while (top > 0) {
        top--;
        int64_t a = st[2 * top], b = st[2 * top + 1];
        const orc_cell *A = &T->cells[a], *B = &T->cells[b];
        if (top + 64 >= cap) {
            cap *= 2;
            st = realloc(st, sizeof(int64_t) * 2 * cap);
        }
        if (a == b) {
            if (A->nchild == 0) {
                pl_push(p2p, a, b);
            } else {
                for (int i = A->nchild - 1; i >= 0; i--)
                    for (int j = A->nchild - 1; j >= 0; j--) {
                        st[2 * top] = A->child0 + i; st[2 * top + 1] = A->child0 + j; top++;
                    }
            }
        } else if (orc_fmm_mac(A->c, A->R, B->c, B->R, B->smax, theta, sigcut)) {
            pl_push(m2l, a, b);
        } else if (A->nchild == 0 && B->nchild == 0) {
            pl_push(p2p, a, b);
        } else if (B->nchild == 0 || (A->nchild != 0 && A->R >= B->R)) {
            for (int i = A->nchild - 1; i >= 0; i--) {
                st[2 * top] = A->child0 + i; st[2 * top + 1] = b; top++;
            }
        } else {
            for (int j = B->nchild - 1; j >= 0; j--) {
                st[2 * top] = a; st[2 * top + 1] = B->child0 + j; top++;
            }
        }
    }
    free(st);
    qsort(m2l->a, m2l->n, 2 * sizeof(int64_t), cmp_pair);
    qsort(p2p->a, p2p->n, 2 * sizeof(int64_t), cmp_pair);
}

/* CSR of a sorted pair list by target */
static int64_t *csr_offsets(const pairlist *L, int64_t ncell)
{
    int64_t *off = calloc(ncell + 1, sizeof(int64_t));
    for (int64_t i = 0; i < L->n; i++) off[L->a[2 * i] + 1]++;
    for (int64_t c = 0; c < ncell; c++) off[c + 1] += off[c];
    return off;
}

/* ---- FMM evaluation --------------------------------------------------------------------- */
typedef struct {
    double t_tree, t_traverse, t_p2m, t_m2m, t_m2l, t_l2l, t_l2p, t_p2p;
    int64_t ncell, nleaf, nm2l, np2p, p2p_pairs, nactive_leaf;
    /* work of the full evaluation vs the evaluated (active) part */
    int64_t p2p_pairs_all, nm2l_active, ncell_active;
} orc_fmm_stats;

static double wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* U (n,3) and J (n,9), overwritten.  X, G AoS (n,3); sig (n).  `sample`
 * (optional, length nsample, sorted target-cell mask built from particle
 * indices) limits P2P/M2L/L2P to the leaves holding those particles, for a
 * bounded-time baseline; pass NULL for the full evaluation. */
int orc_uj_fmm(int kernel, int p, int ncrit, double theta, double sigcut, int64_t n,
               const double *X, const double *G, const double *sig, double *U, double *J,
               const int64_t *sample, int64_t nsample, orc_fmm_stats *stats)
{
    if (p < 0 || p > MAXP - 2 || ncrit < 1) return 1;
    orc_fmm_stats st;
    memset(&st, 0, sizeof(st));
    double t0 = wtime();
    orc_tree T;
    tree_build(&T, n, X, sig, ncrit);
    double t1 = wtime();
    st.t_tree = t1 - t0;
    pairlist m2l = {0}, p2p = {0};
    if (n > 0) traverse(&T, theta, sigcut, &m2l, &p2p);
    double t2 = wtime();
    st.t_traverse = t2 - t1;
    int64_t nc = T.ncell;
    int64_t *m2l_off = csr_offsets(&m2l, nc), *p2p_off = csr_offsets(&p2p, nc);

    /* which cells are evaluated (targets) */
    char *active = calloc(nc > 0 ? nc : 1, 1);
    int64_t *leaf_of = malloc(sizeof(int64_t) * (n > 0 ? n : 1));   /* sorted slot -> leaf */
    for (int64_t c = 0; c < nc; c++)
        if (T.cells[c].nchild == 0)
            for (int64_t j = T.cells[c].begin; j < T.cells[c].end; j++) leaf_of[j] = c;
    if (!sample) {
        memset(active, 1, nc);
    } else {
        int64_t *slot = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
        for (int64_t j = 0; j < n; j++) slot[T.perm[j]] = j;
        for (int64_t s = 0; s < nsample; s++) {
            int64_t c = leaf_of[slot[sample[s]]];
            active[c] = 1;
        }
        free(slot);
        /* ancestors of active leaves need their locals */
        for (int64_t c = nc - 1; c >= 0; c--)
            if (T.cells[c].nchild)
                for (int k = 0; k < T.cells[c].nchild; k++)
                    if (active[T.cells[c].child0 + k]) active[c] = 1;
    }

    mindex M;
    mi_build(&M, p);
    const int K = M.K;
    double *mp = calloc((size_t)nc * 3 * K, sizeof(double));   /* multipoles m[c][comp][K] */
    double *lc = calloc((size_t)nc * 3 * K, sizeof(double));   /* locals */

    /* P2M */
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < nc; c++) {
        const orc_cell *C = &T.cells[c];
        if (C->nchild) continue;
        double mono[1024];
        for (int64_t j = C->begin; j < C->end; j++) {
            int64_t i = T.perm[j];
            double t[3] = {X[3 * i] - C->c[0], X[3 * i + 1] - C->c[1], X[3 * i + 2] - C->c[2]};
            monomials(&M, t, mono);
            for (int q = 0; q < 3; q++) {
                double g = G[3 * i + q];
                double *m = mp + ((size_t)c * 3 + q) * K;
                for (int k = 0; k < K; k++) m[k] += g * mono[k];
            }
        }
    }
    double t3 = wtime();
    st.t_p2m = t3 - t2;

    /* M2M, deepest level first */
    int maxlev = 0;
    for (int64_t c = 0; c < nc; c++) maxlev = T.cells[c].level > maxlev ? T.cells[c].level : maxlev;
    for (int lev = maxlev - 1; lev >= 0; lev--) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t c = 0; c < nc; c++) {
            const orc_cell *C = &T.cells[c];
            if (C->level != lev || C->nchild == 0) continue;
            double dm[1024];
            for (int ch = 0; ch < C->nchild; ch++) {
                const orc_cell *D = &T.cells[C->child0 + ch];
                double d[3] = {D->c[0] - C->c[0], D->c[1] - C->c[1], D->c[2] - C->c[2]};
                monomials(&M, d, dm);
                for (int q = 0; q < 3; q++) {
                    const double *mc = mp + ((size_t)(C->child0 + ch) * 3 + q) * K;
                    double *m = mp + ((size_t)c * 3 + q) * K;
                    for (int k = 0; k < K; k++) {
                        const int *a = M.mi[k];
                        double s = 0.0;
                        for (int b0 = 0; b0 <= a[0]; b0++)
                            for (int b1 = 0; b1 <= a[1]; b1++)
                                for (int b2 = 0; b2 <= a[2]; b2++) {
                                    int jb = mi_idx(&M, b0, b1, b2);
                                    int jd = mi_idx(&M, a[0] - b0, a[1] - b1, a[2] - b2);
                                    s += binom(a[0], b0) * binom(a[1], b1) * binom(a[2], b2) *
                                         mc[jb] * dm[jd];
                                }
                        m[k] += s;
                    }
                }
            }
        }
    }
    double t4 = wtime();
    st.t_m2m = t4 - t3;

    /* M2L table: (a, b, a+b, sign * C(a+b, a)) with |a| + |b| <= p */
    int64_t ntab = 0;
    for (int ka = 0; ka < K; ka++)
        for (int kb = 0; kb < K; kb++) {
            const int *a = M.mi[ka], *b = M.mi[kb];
            if (a[0] + a[1] + a[2] + b[0] + b[1] + b[2] <= p) ntab++;
        }
    int *tab = malloc(sizeof(int) * 3 * (ntab > 0 ? ntab : 1));
    double *tcoef = malloc(sizeof(double) * (ntab > 0 ? ntab : 1));
    ntab = 0;
    for (int kb = 0; kb < K; kb++)
        for (int ka = 0; ka < K; ka++) {
            const int *a = M.mi[ka], *b = M.mi[kb];
            int na = a[0] + a[1] + a[2];
            if (na + b[0] + b[1] + b[2] > p) continue;
            tab[3 * ntab] = ka;
            tab[3 * ntab + 1] = kb;
            tab[3 * ntab + 2] = mi_idx(&M, a[0] + b[0], a[1] + b[1], a[2] + b[2]);
            tcoef[ntab] = (na % 2 ? -1.0 : 1.0) * binom(a[0] + b[0], a[0]) *
                          binom(a[1] + b[1], a[1]) * binom(a[2] + b[2], a[2]);
            ntab++;
        }

    /* M2L */
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t c = 0; c < nc; c++) {
        if (!active[c]) continue;
        const orc_cell *A = &T.cells[c];
        double bb[1024];
        for (int64_t e = m2l_off[c]; e < m2l_off[c + 1]; e++) {
            int64_t s = m2l.a[2 * e + 1];
            const orc_cell *B = &T.cells[s];
            double r[3] = {A->c[0] - B->c[0], A->c[1] - B->c[1], A->c[2] - B->c[2]};
            deriv_b(&M, r, bb);
            for (int q = 0; q < 3; q++) {
                const double *m = mp + ((size_t)s * 3 + q) * K;
                double *l = lc + ((size_t)c * 3 + q) * K;
                for (int64_t t = 0; t < ntab; t++)
                    l[tab[3 * t + 1]] += tcoef[t] * m[tab[3 * t]] * bb[tab[3 * t + 2]];
            }
        }
    }
    double t5 = wtime();
    st.t_m2l = t5 - t4;

    /* L2L, top-down */
    for (int lev = 1; lev <= maxlev; lev++) {
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t c = 0; c < nc; c++) {
            const orc_cell *C = &T.cells[c];
            if (C->level != lev - 1 || C->nchild == 0) continue;
            for (int ch = 0; ch < C->nchild; ch++) {
                int64_t d = C->child0 + ch;
                if (!active[d]) continue;
                const orc_cell *D = &T.cells[d];
                double e[3] = {D->c[0] - C->c[0], D->c[1] - C->c[1], D->c[2] - C->c[2]};
                double em[1024];
                monomials(&M, e, em);
                for (int q = 0; q < 3; q++) {
                    const double *lp = lc + ((size_t)c * 3 + q) * K;
                    double *l = lc + ((size_t)d * 3 + q) * K;
                    for (int kg = 0; kg < K; kg++) {
                        const int *g = M.mi[kg];
                        double s = 0.0;
                        for (int kb = 0; kb < K; kb++) {
                            const int *b = M.mi[kb];
                            if (b[0] < g[0] || b[1] < g[1] || b[2] < g[2]) continue;
                            int je = mi_idx(&M, b[0] - g[0], b[1] - g[1], b[2] - g[2]);
                            s += lp[kb] * binom(b[0], g[0]) * binom(b[1], g[1]) * binom(b[2], g[2]) *
                                 em[je];
                        }
                        l[kg] += s;
                    }
                }
            }
        }
    }
    double t6 = wtime();
    st.t_l2l = t6 - t5;

    /* L2P: far-field U, J from the leaf's local expansion of phi_q */
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < nc; c++) {
        const orc_cell *C = &T.cells[c];
        if (C->nchild || !active[c]) continue;
        double mono[1024];
        for (int64_t j = C->begin; j < C->end; j++) {
            int64_t i = T.perm[j];
            double t[3] = {X[3 * i] - C->c[0], X[3 * i + 1] - C->c[1], X[3 * i + 2] - C->c[2]};
            monomials(&M, t, mono);
            double d1[3][3] = {{0}}, d2[3][3][3] = {{{0}}};   /* d1[q][j], d2[q][j][l] */
            for (int kb = 0; kb < K; kb++) {
                const int *b = M.mi[kb];
                for (int jj = 0; jj < 3; jj++) {
                    int e1[3] = {b[0], b[1], b[2]};
                    e1[jj]++;
                    int k1 = mi_idx(&M, e1[0], e1[1], e1[2]);
                    if (k1 < 0) continue;
                    double f1 = (double)e1[jj] * mono[kb];
                    for (int q = 0; q < 3; q++) d1[q][jj] += lc[((size_t)c * 3 + q) * K + k1] * f1;
                    for (int ll = jj; ll < 3; ll++) {
                        int e2[3] = {e1[0], e1[1], e1[2]};
                        e2[ll]++;
                        int k2 = mi_idx(&M, e2[0], e2[1], e2[2]);
                        if (k2 < 0) continue;
                        double f2 = f1 * (double)e2[ll];
                        for (int q = 0; q < 3; q++) d2[q][jj][ll] += lc[((size_t)c * 3 + q) * K + k2] * f2;
                    }
                }
            }
            for (int q = 0; q < 3; q++)
                for (int jj = 0; jj < 3; jj++)
                    for (int ll = 0; ll < jj; ll++) d2[q][jj][ll] = d2[q][ll][jj];
            /* u_i = (1/4pi) eps_ijk d_j phi_k ; J_il = (1/4pi) eps_ijk d_l d_j phi_k */
            double u[3], Jm[9];
            u[0] = d1[2][1] - d1[1][2];
            u[1] = d1[0][2] - d1[2][0];
            u[2] = d1[1][0] - d1[0][1];
            for (int l = 0; l < 3; l++) {
                Jm[0 + 3 * l] = d2[2][1][l] - d2[1][2][l];
                Jm[1 + 3 * l] = d2[0][2][l] - d2[2][0][l];
                Jm[2 + 3 * l] = d2[1][0][l] - d2[0][1][l];
            }
            for (int k = 0; k < 3; k++) U[3 * i + k] = ORC_CONST4 * u[k];
            for (int k = 0; k < 9; k++) J[9 * i + k] = ORC_CONST4 * Jm[k];
        }
    }
    double t7 = wtime();
    st.t_l2p = t7 - t6;

    /* P2P (regularised, direct) */
    int64_t pairs = 0;
#pragma omp parallel for schedule(dynamic, 2) reduction(+ : pairs)
    for (int64_t c = 0; c < nc; c++) {
        const orc_cell *A = &T.cells[c];
        if (A->nchild || !active[c]) continue;
        for (int64_t jt = A->begin; jt < A->end; jt++) {
            int64_t i = T.perm[jt];
            double u[3] = {0, 0, 0}, jj[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t e = p2p_off[c]; e < p2p_off[c + 1]; e++) {
                const orc_cell *B = &T.cells[p2p.a[2 * e + 1]];
                for (int64_t js = B->begin; js < B->end; js++) {
                    int64_t s = T.perm[js];
                    orc_pair_uj(kernel, X + 3 * i, X + 3 * s, G + 3 * s, sig[s], u, jj);
                }
                pairs += B->end - B->begin;
            }
            for (int k = 0; k < 3; k++) U[3 * i + k] += u[k];
            for (int k = 0; k < 9; k++) J[9 * i + k] += jj[k];
        }
    }
    double t8 = wtime();
    st.t_p2p = t8 - t7;

    st.ncell = nc;
    for (int64_t c = 0; c < nc; c++) {
        st.nleaf += T.cells[c].nchild == 0;
        st.nactive_leaf += T.cells[c].nchild == 0 && active[c];
        st.ncell_active += active[c] != 0;
        if (active[c]) st.nm2l_active += m2l_off[c + 1] - m2l_off[c];
        if (T.cells[c].nchild == 0) {
            int64_t s = 0;
            for (int64_t e = p2p_off[c]; e < p2p_off[c + 1]; e++)
                s += T.cells[p2p.a[2 * e + 1]].end - T.cells[p2p.a[2 * e + 1]].begin;
            st.p2p_pairs_all += s * (T.cells[c].end - T.cells[c].begin);
        }
    }
    st.nm2l = m2l.n;
    st.np2p = p2p.n;
    st.p2p_pairs = pairs;
    if (stats) *stats = st;

    free(tab); free(tcoef); free(mp); free(lc); mi_free(&M);
    free(active); free(leaf_of); free(m2l_off); free(p2p_off);
    free(m2l.a); free(p2p.a);
    tree_free(&T);
    return 0;
}

void orc_pair_dj(int kernel, const double *xt, const double *xs, const double *gs, double sgm,
                 double *DU, double *DJ);
void orc_pair_estr(int kernel, int transposed, const double *xt, const double *tJ,
                   const double *tDJ, const double *xs, const double *gs, double sg,
                   const double *sJ, const double *sDJ, double *e, double *de);

/* Sums over the near-field (P2P) lists of the tree and traversal orc_uj_fmm
 * builds for the same particles, ncrit, theta and sigcut -- the part of a
 * sum that the FMM evaluates pair by pair.  Outputs are overwritten.
 *   what = 1: core-size derivatives D[U] (out1, n x 3) and D[J] (out2, n x 9)
 *             of orc_pair_dj.  Their far field is zero: the multipole part of
 *             the FMM uses the singular kernel, which has no core.
 *   what = 2: Estr (out1, n x 3) of orc_pair_estr from J (n x 9), and with DJ
 *             (n x 9) also D[E] (out2, n x 3).  The SFS model sums zeta, which
 *             the FMM leaves to its near field as well.
 * Each target sums its list's source leaves in list order, each leaf's
 * particles in sorted order. */
int orc_fmm_near(int kernel, int transposed, int what, int ncrit, double theta, double sigcut,
                 int64_t n, const double *X, const double *G, const double *sig, const double *J,
                 const double *DJ, double *out1, double *out2)
{
    if (ncrit < 1 || (what != 1 && what != 2) || (what == 2 && !J)) return 1;
    const int w1 = 3, w2 = what == 1 ? 9 : 3;
    memset(out1, 0, sizeof(double) * w1 * (size_t)n);
    if (what == 1 || DJ) memset(out2, 0, sizeof(double) * w2 * (size_t)n);
    if (n == 0) return 0;
    orc_tree T;
    tree_build(&T, n, X, sig, ncrit);
    pairlist m2l = {0}, p2p = {0};
    traverse(&T, theta, sigcut, &m2l, &p2p);
    const int64_t nc = T.ncell;
    int64_t *p2p_off = csr_offsets(&p2p, nc);
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t c = 0; c < nc; c++) {
        const orc_cell *A = &T.cells[c];
        if (A->nchild) continue;
        for (int64_t jt = A->begin; jt < A->end; jt++) {
            const int64_t i = T.perm[jt];
            double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t e = p2p_off[c]; e < p2p_off[c + 1]; e++) {
                const orc_cell *B = &T.cells[p2p.a[2 * e + 1]];
                for (int64_t js = B->begin; js < B->end; js++) {
                    const int64_t s = T.perm[js];
                    if (what == 1)
                        orc_pair_dj(kernel, X + 3 * i, X + 3 * s, G + 3 * s, sig[s], a, b);
                    else
                        orc_pair_estr(kernel, transposed, X + 3 * i, J + 9 * i,
                                      DJ ? DJ + 9 * i : NULL, X + 3 * s, G + 3 * s, sig[s],
                                      J + 9 * s, DJ ? DJ + 9 * s : NULL, a, b);
                }
            }
            for (int k = 0; k < w1; k++) out1[w1 * i + k] = a[k];
            if (what == 1 || DJ)
                for (int k = 0; k < w2; k++) out2[w2 * i + k] = b[k];
        }
    }
    free(p2p_off);
    free(m2l.a); free(p2p.a);
    tree_free(&T);
    return 0;
}

/* ---- viscous sums over the near-field lists ------------------------------------
 * The FMM leaves the vorticity kernel zeta (and PSE's eta = -zeta'/rho) to its
 * near field, as its rbf mode does for FLOWVPM's zeta_fmm
 * (examples/rotorhover/rotorhover.jl:170): a plan is the tree and P2P lists
 * orc_uj_fmm builds for the same particles, sigmas, ncrit, theta and sigcut,
 * and each sum runs over a target leaf's list in list order, each source
 * leaf's particles in sorted order. */
struct orc_near_plan {
    orc_tree T;
    pairlist m2l, p2p;
    int64_t *p2p_off;
};

struct orc_near_plan *orc_near_plan_build(int ncrit, double theta, double sigcut, int64_t n,
                                          const double *X, const double *sig)
{
    if (ncrit < 1 || n < 0) return NULL;
    struct orc_near_plan *pl = calloc(1, sizeof(*pl));
    tree_build(&pl->T, n, X, sig, ncrit);
    traverse(&pl->T, theta, sigcut, &pl->m2l, &pl->p2p);
    pl->p2p_off = csr_offsets(&pl->p2p, pl->T.ncell);
    return pl;
}

void orc_near_plan_free(struct orc_near_plan *pl)
{
    if (!pl) return;
    free(pl->p2p_off);
    free(pl->m2l.a);
    free(pl->p2p.a);
    tree_free(&pl->T);
    free(pl);
}

double orc_eta(int kernel, double rho);
void orc_kernel(int kernel, double rho, double *zeta, double *g, double *dg);

/* w_i = sum_j Gamma_j zeta(r_ij / s_j) / s_j^3 over the near lists (s_j =
 * ssig[j], or s0 for every source when s0 > 0; masked sources have no
 * strength; coincident points included), w (n x 3) overwritten. */
void orc_near_zeta(const struct orc_near_plan *pl, int kernel, const double *X, const double *G,
                   const double *ssig, double s0, const uint8_t *mask, double *w)
{
    const orc_tree *T = &pl->T;
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t c = 0; c < T->ncell; c++) {
        const orc_cell *A = &T->cells[c];
        if (A->nchild) continue;
        for (int64_t jt = A->begin; jt < A->end; jt++) {
            const int64_t i = T->perm[jt];
            double acc[3] = {0, 0, 0};
            for (int64_t e = pl->p2p_off[c]; e < pl->p2p_off[c + 1]; e++) {
                const orc_cell *B = &T->cells[pl->p2p.a[2 * e + 1]];
                for (int64_t js = B->begin; js < B->end; js++) {
                    const int64_t j = T->perm[js];
                    if (mask && mask[j]) continue;
                    const double d0 = X[3 * i] - X[3 * j], d1 = X[3 * i + 1] - X[3 * j + 1],
                                 d2 = X[3 * i + 2] - X[3 * j + 2];
                    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    const double sj = s0 > 0.0 ? s0 : ssig[j];
                    double z;
                    orc_kernel(kernel, r / sj, &z, NULL, NULL);
                    z /= sj * sj * sj;
                    for (int k = 0; k < 3; k++) acc[k] += z * G[3 * j + k];
                }
            }
            for (int k = 0; k < 3; k++) w[3 * i + k] = acc[k];
        }
    }
}

/* PSE dGamma/dt (orc_pse's pair terms) over the near lists, V (n x 3)
 * overwritten. */
void orc_near_pse(const struct orc_near_plan *pl, int kernel, double nu, const double *X,
                  const double *G, const double *sig, const double *vol, double *V)
{
    const orc_tree *T = &pl->T;
#pragma omp parallel for schedule(dynamic, 2)
    for (int64_t c = 0; c < T->ncell; c++) {
        const orc_cell *A = &T->cells[c];
        if (A->nchild) continue;
        for (int64_t jt = A->begin; jt < A->end; jt++) {
            const int64_t i = T->perm[jt];
            double acc[3] = {0, 0, 0};
            for (int64_t e = pl->p2p_off[c]; e < pl->p2p_off[c + 1]; e++) {
                const orc_cell *B = &T->cells[pl->p2p.a[2 * e + 1]];
                for (int64_t js = B->begin; js < B->end; js++) {
                    const int64_t j = T->perm[js];
                    if (j == i) continue;
                    const double d0 = X[3 * i] - X[3 * j], d1 = X[3 * i + 1] - X[3 * j + 1],
                                 d2 = X[3 * i + 2] - X[3 * j + 2];
                    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    const double s2 = 0.5 * (sig[i] * sig[i] + sig[j] * sig[j]);
                    const double s = sqrt(s2);
                    const double wgt = 2.0 * nu / s2 * orc_eta(kernel, r / s) / (s2 * s);
                    for (int k = 0; k < 3; k++)
                        acc[k] += wgt * (vol[i] * G[3 * j + k] - vol[j] * G[3 * i + k]);
                }
            }
            for (int k = 0; k < 3; k++) V[3 * i + k] = acc[k];
        }
    }
}

/* Expose the derivative recurrence for tests: b_k(r), |k| <= P, graded order. */
int orc_fmm_deriv(int P, const double r[3], double *b, int *mi_out)
{
    if (P < 0 || P > MAXP) return -1;
    mindex M;
    mi_build(&M, P);
    deriv_b(&M, r, b);
    if (mi_out) memcpy(mi_out, M.mi, sizeof(int[3]) * M.K);
    int K = M.K;
    mi_free(&M);
    return K;
}
