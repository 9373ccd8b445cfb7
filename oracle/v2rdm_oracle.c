/*
 * v2rdm_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see header).
 *
 * Plain-loop restatement of edeprince3/hilbert's v2RDM-CASSCF setup and
 * DQG+spin constraint maps.  Au and ATu are written as separate loops, as in
 * the reference, so that the adjointness test checks one against the other.
 */
#define _GNU_SOURCE
#include "v2rdm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- helpers ------------------------------------------------------------ */
#define NMO      (p->nmo)
#define SYM(i)   (p->symmetry[(i)])
#define GAB(h)   (p->gems_ab[(h)])
#define GAA(h)   (p->gems_aa[(h)])
#define IAB(h, i, j) (p->ibas_ab[(h)][(i) * NMO + (j)])
#define IAA(h, i, j) (p->ibas_aa[(h)][(i) * NMO + (j)])
#define BAB0(h, n) (p->bas_ab[(h)][2 * (n)])
#define BAB1(h, n) (p->bas_ab[(h)][2 * (n) + 1])
#define BAA0(h, n) (p->bas_aa[(h)][2 * (n)])
#define BAA1(h, n) (p->bas_aa[(h)][2 * (n) + 1])
/* element (i,j) of a one-body block, i and j absolute orbital indices of one irrep */
#define OB(off, i, j) ((off)[SYM(i)] + (long)((i) - p->pitzer_offset[SYM(i)]) * p->amopi[SYM(i)] \
                       + ((j) - p->pitzer_offset[SYM(i)]))
#define SGN(a, b) ((a) < (b) ? 1.0 : -1.0)

static double wall(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* ---- setup -------------------------------------------------------------- */

/* basis.cc:47-76 (product table = XOR of Cotton indices), 104-111 (orbital
 * irreps), 413-419 (pitzer_offset), 427-533 (geminal tables). */
static void build_tables(orc_problem *p) {
    int n = p->nmo;
    memset(p->amopi, 0, sizeof(p->amopi));
    for (int i = 0; i < n; i++) p->amopi[p->symmetry[i]]++;
    int acc = 0;
    for (int h = 0; h < p->nirrep; h++) { p->pitzer_offset[h] = acc; acc += p->amopi[h]; }
    for (int h = 0; h < p->nirrep; h++) {
        p->ibas_ab[h] = (int *)malloc(sizeof(int) * n * n);
        p->ibas_aa[h] = (int *)malloc(sizeof(int) * n * n);
        p->bas_ab[h] = (int *)malloc(sizeof(int) * 2 * (n * n + 1));
        p->bas_aa[h] = (int *)malloc(sizeof(int) * 2 * (n * n + 1));
        for (int t = 0; t < n * n; t++) { p->ibas_ab[h][t] = -999; p->ibas_aa[h][t] = -999; }
        for (int t = 0; t < 2 * (n * n + 1); t++) { p->bas_ab[h][t] = -999; p->bas_aa[h][t] = -999; }
        int nab = 0, naa = 0;
        /* outer loop over the second orbital, inner over the first: the pair
         * pushed is (first = inner, second = outer) */
        for (int outer = 0; outer < n; outer++) {
            for (int inner = 0; inner < n; inner++) {
                if ((p->symmetry[outer] ^ p->symmetry[inner]) != h) continue;
                int a = inner, b = outer;
                p->ibas_ab[h][a * n + b] = nab;
                p->bas_ab[h][2 * nab] = a;
                p->bas_ab[h][2 * nab + 1] = b;
                nab++;
                if (a < b) {
                    p->ibas_aa[h][a * n + b] = naa;
                    p->ibas_aa[h][b * n + a] = naa;
                    p->bas_aa[h][2 * naa] = a;
                    p->bas_aa[h][2 * naa + 1] = b;
                    naa++;
                }
            }
        }
        p->gems_ab[h] = nab;
        p->gems_aa[h] = naa;
    }
}

/* v2rdm_solver.cc:3294-3450 (n_primal), 3452-3727 (n_dual), 3729-3846 (offsets). */
static void build_offsets(orc_problem *p) {
    int H = p->nirrep;
    long off = 0;
    int nb = 0;
#define PUSH(arr, dim) do { (arr)[h] = off; off += (long)(dim) * (dim); p->dims[nb++] = (dim); } while (0)
    for (int h = 0; h < H; h++) PUSH(p->d2aboff, GAB(h));
    for (int h = 0; h < H; h++) PUSH(p->d2aaoff, GAA(h));
    for (int h = 0; h < H; h++) PUSH(p->d2bboff, GAA(h));
    if (p->constrain_spin) {
        int singlet = (p->nalpha == p->nbeta);
        for (int h = 0; h < H; h++) PUSH(p->d200off, singlet ? GAB(h) : 2 * GAB(h));
    }
    for (int h = 0; h < H; h++) PUSH(p->d1aoff, p->amopi[h]);
    for (int h = 0; h < H; h++) PUSH(p->d1boff, p->amopi[h]);
    for (int h = 0; h < H; h++) PUSH(p->q1aoff, p->amopi[h]);
    for (int h = 0; h < H; h++) PUSH(p->q1boff, p->amopi[h]);
    if (p->constrain_q2) {
        for (int h = 0; h < H; h++) PUSH(p->q2aboff, GAB(h));
        for (int h = 0; h < H; h++) PUSH(p->q2aaoff, GAA(h));
        for (int h = 0; h < H; h++) PUSH(p->q2bboff, GAA(h));
    }
    if (p->constrain_g2) {
        for (int h = 0; h < H; h++) PUSH(p->g2aboff, GAB(h));
        for (int h = 0; h < H; h++) PUSH(p->g2baoff, GAB(h));
        for (int h = 0; h < H; h++) PUSH(p->g2aaoff, 2 * GAB(h));
    }
#undef PUSH
    p->n_primal = off;
    p->nblocks = nb;

    long nd = p->constrain_sz ? 3 : 1;
    long sab = 0, saa = 0, sm = 0;
    for (int h = 0; h < H; h++) {
        sab += (long)GAB(h) * GAB(h);
        saa += (long)GAA(h) * GAA(h);
        sm += (long)p->amopi[h] * p->amopi[h];
    }
    nd += 2 * saa + sab;                /* hermiticity */
    nd += 2 * sm;                       /* D1 <-> Q1 */
    nd += (p->constrain_sz ? 4 : 2) * sm; /* contractions */
    if (p->constrain_spin) {
        nd += 1;
        if (p->nalpha == p->nbeta) nd += sm + 3 * saa + 2 * sab;
        else nd += 4 * sab;
        if (p->constrain_g2) nd += 2 * GAB(0);
    }
    if (p->constrain_q2) nd += sab + 2 * saa;
    if (p->constrain_g2) nd += 2 * sab + 4 * sab;
    p->n_dual = nd;
}

int orc_problem_init(orc_problem *p, int nmo, int nalpha, int nbeta, int multiplicity,
                     const int *irreps, int nirrep,
                     int q2, int g2, int spin, int sz) {
    memset(p, 0, sizeof(*p));
    if (nmo <= 0 || nirrep < 1 || nirrep > ORC_MAX_IRREP) return -1;
    if (nalpha < 0 || nbeta < 0 || nalpha > nmo || nbeta > nmo) return -2;
    p->nmo = nmo;
    p->nirrep = nirrep;
    p->nalpha = nalpha;
    p->nbeta = nbeta;
    p->multiplicity = multiplicity;
    p->constrain_q2 = q2;
    p->constrain_g2 = g2;
    p->constrain_spin = spin;
    p->constrain_sz = sz;
    p->symmetry = (int *)malloc(sizeof(int) * nmo);
    for (int i = 0; i < nmo; i++) {
        if (irreps[i] < 0 || irreps[i] >= nirrep) return -3;
        if (i > 0 && irreps[i] < irreps[i - 1]) return -4; /* Pitzer order required */
        p->symmetry[i] = irreps[i];
    }
    build_tables(p);
    build_offsets(p);
    return 0;
}

void orc_problem_free(orc_problem *p) {
    for (int h = 0; h < p->nirrep; h++) {
        free(p->ibas_ab[h]); free(p->ibas_aa[h]);
        free(p->bas_ab[h]); free(p->bas_aa[h]);
        free(p->ibas_aab[h]); free(p->ibas_aaa[h]);
        free(p->bas_aab[h]); free(p->bas_aaa[h]); free(p->bas_aba[h]);
    }
    free(p->symmetry);
    memset(p, 0, sizeof(*p));
}

/* basis.cc:642-744: triplets (i,j,k) of irrep h in loop order i, j, k; aba
 * keeps all, aab keeps i < j, aaa keeps i < j < k; the index tables are
 * filled for the swapped first two (aab) and for all six orders (aaa). */
static void build_triplets(orc_problem *p) {
    const int n = NMO;
    const long n3 = (long)n * n * n;
    for (int h = 0; h < p->nirrep; h++) {
        p->ibas_aab[h] = (int *)malloc(sizeof(int) * n3);
        p->ibas_aaa[h] = (int *)malloc(sizeof(int) * n3);
        p->bas_aab[h] = (int *)malloc(sizeof(int) * 3 * n3);
        p->bas_aaa[h] = (int *)malloc(sizeof(int) * 3 * n3);
        p->bas_aba[h] = (int *)malloc(sizeof(int) * 3 * n3);
        for (long e = 0; e < n3; e++) p->ibas_aab[h][e] = p->ibas_aaa[h][e] = -999;
        int caab = 0, caaa = 0, caba = 0;
#define I3(a, b, c) (((long)(a) * n + (b)) * n + (c))
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int k = 0; k < n; k++) {
                    if ((SYM(i) ^ SYM(j) ^ SYM(k)) != h) continue;
                    p->bas_aba[h][3 * caba] = i;
                    p->bas_aba[h][3 * caba + 1] = j;
                    p->bas_aba[h][3 * caba + 2] = k;
                    caba++;
                    if (i >= j) continue;
                    p->ibas_aab[h][I3(i, j, k)] = p->ibas_aab[h][I3(j, i, k)] = caab;
                    p->bas_aab[h][3 * caab] = i;
                    p->bas_aab[h][3 * caab + 1] = j;
                    p->bas_aab[h][3 * caab + 2] = k;
                    caab++;
                    if (j >= k) continue;
                    p->ibas_aaa[h][I3(i, j, k)] = p->ibas_aaa[h][I3(i, k, j)] = p->ibas_aaa[h][I3(j, i, k)] = caaa;
                    p->ibas_aaa[h][I3(j, k, i)] = p->ibas_aaa[h][I3(k, i, j)] = p->ibas_aaa[h][I3(k, j, i)] = caaa;
                    p->bas_aaa[h][3 * caaa] = i;
                    p->bas_aaa[h][3 * caaa + 1] = j;
                    p->bas_aaa[h][3 * caaa + 2] = k;
                    caaa++;
                }
#undef I3
        p->trip_aab[h] = caab;
        p->trip_aaa[h] = caaa;
        p->trip_aba[h] = caba;
    }
}

int orc_problem_enable_t(orc_problem *p, int t1, int t2, int d3) {
    if ((t1 || t2) && (!p->constrain_q2 || !p->constrain_g2)) return -1;
    if (p->constrain_t1 || p->constrain_t2 || p->constrain_d3) return -1;
    if (!t1 && !t2 && !d3) return 0;
    p->constrain_t1 = t1 != 0;
    p->constrain_t2 = t2 != 0;
    p->constrain_d3 = d3 != 0;
    build_triplets(p);
    const int H = p->nirrep;
    long off = p->n_primal;
    int nb = p->nblocks;
#define PUSHT(arr, dim) do { (arr)[h] = off; off += (long)(dim) * (dim); p->dims[nb++] = (dim); \
                             p->n_dual += (long)(dim) * (dim); } while (0)
    if (t1) { /* v2rdm_solver.cc:3848-3873; rows :3554-3567 */
        for (int h = 0; h < H; h++) PUSHT(p->t1aaaoff, p->trip_aaa[h]);
        for (int h = 0; h < H; h++) PUSHT(p->t1bbboff, p->trip_aaa[h]);
        for (int h = 0; h < H; h++) PUSHT(p->t1aaboff, p->trip_aab[h]);
        for (int h = 0; h < H; h++) PUSHT(p->t1bbaoff, p->trip_aab[h]);
    }
    if (t2) { /* v2rdm_solver.cc:3879-3908; rows :3568-3581 */
        for (int h = 0; h < H; h++) PUSHT(p->t2aaaoff, p->trip_aab[h] + p->trip_aba[h]);
        for (int h = 0; h < H; h++) PUSHT(p->t2bbboff, p->trip_aab[h] + p->trip_aba[h]);
        for (int h = 0; h < H; h++) PUSHT(p->t2aaboff, p->trip_aab[h]);
        for (int h = 0; h < H; h++) PUSHT(p->t2bbaoff, p->trip_aab[h]);
    }
#undef PUSHT
    if (d3) { /* v2rdm_solver.cc:4000-4026; rows :3624-3675 */
        for (int h = 0; h < H; h++) { p->d3aaaoff[h] = off; off += (long)p->trip_aaa[h] * p->trip_aaa[h]; p->dims[nb++] = p->trip_aaa[h]; }
        for (int h = 0; h < H; h++) { p->d3bbboff[h] = off; off += (long)p->trip_aaa[h] * p->trip_aaa[h]; p->dims[nb++] = p->trip_aaa[h]; }
        for (int h = 0; h < H; h++) { p->d3aaboff[h] = off; off += (long)p->trip_aab[h] * p->trip_aab[h]; p->dims[nb++] = p->trip_aab[h]; }
        for (int h = 0; h < H; h++) { p->d3bbaoff[h] = off; off += (long)p->trip_aab[h] * p->trip_aab[h]; p->dims[nb++] = p->trip_aab[h]; }
        for (int h = 0; h < H; h++) {
            const long saa = (long)GAA(h) * GAA(h), sab = (long)GAB(h) * GAB(h);
            p->n_dual += p->constrain_sz ? 4 * saa + 2 * sab : 2 * saa + sab;
            if (p->constrain_spin && p->nalpha == p->nbeta)
                p->n_dual += (long)p->trip_aab[h] * p->trip_aab[h] + 2L * p->trip_aaa[h] * p->trip_aaa[h];
        }
    }
    p->n_primal = off;
    p->nblocks = nb;
    return 0;
}

void orc_get_t_offsets(const orc_problem *p, long *off) {
    const long *arrs[12] = {p->t1aaaoff, p->t1bbboff, p->t1aaboff, p->t1bbaoff,
                            p->t2aaaoff, p->t2bbboff, p->t2aaboff, p->t2bbaoff,
                            p->d3aaaoff, p->d3bbboff, p->d3aaboff, p->d3bbaoff};
    for (int a = 0; a < 12; a++)
        for (int h = 0; h < p->nirrep; h++) off[a * p->nirrep + h] = arrs[a][h];
}

void orc_get_trip_counts(const orc_problem *p, int *taab, int *taaa, int *taba) {
    for (int h = 0; h < p->nirrep; h++) {
        taab[h] = p->trip_aab[h];
        taaa[h] = p->trip_aaa[h];
        taba[h] = p->trip_aba[h];
    }
}

long orc_n_primal(const orc_problem *p) { return p->n_primal; }
long orc_n_dual(const orc_problem *p) { return p->n_dual; }
int orc_nblocks(const orc_problem *p) { return p->nblocks; }
void orc_get_dims(const orc_problem *p, int *dims) { memcpy(dims, p->dims, sizeof(int) * p->nblocks); }
void orc_get_gems(const orc_problem *p, int *gab, int *gaa) {
    for (int h = 0; h < p->nirrep; h++) { gab[h] = GAB(h); gaa[h] = GAA(h); }
}
void orc_get_tables(const orc_problem *p, int *bab, int *iab, int *baa, int *iaa) {
    int n = NMO;
    long ob = 0, oa = 0;
    for (int h = 0; h < p->nirrep; h++) {
        memcpy(bab + ob, p->bas_ab[h], sizeof(int) * 2 * GAB(h)); ob += 2 * GAB(h);
        memcpy(baa + oa, p->bas_aa[h], sizeof(int) * 2 * GAA(h)); oa += 2 * GAA(h);
        memcpy(iab + (long)h * n * n, p->ibas_ab[h], sizeof(int) * n * n);
        memcpy(iaa + (long)h * n * n, p->ibas_aa[h], sizeof(int) * n * n);
    }
}
void orc_get_offsets(const orc_problem *p, long *off) {
    const long *arrs[14] = {p->d2aboff, p->d2aaoff, p->d2bboff, p->d200off, p->d1aoff, p->d1boff,
                            p->q1aoff, p->q1boff, p->q2aboff, p->q2aaoff, p->q2bboff,
                            p->g2aboff, p->g2baoff, p->g2aaoff};
    for (int a = 0; a < 14; a++)
        for (int h = 0; h < p->nirrep; h++) off[a * p->nirrep + h] = arrs[a][h];
}

/* ---- b ------------------------------------------------------------------ */
/* v2rdm_solver.cc:1826-2074 */
void orc_build_b(const orc_problem *p, double *b) {
    memset(b, 0, sizeof(double) * p->n_dual);
    int H = p->nirrep;
    double na = p->nalpha, nb = p->nbeta;
    long r = 0;
    if (p->constrain_sz) {
        b[r++] = na * nb;
        b[r++] = na * (na - 1.0);
        b[r++] = nb * (nb - 1.0);
    } else {
        b[r++] = (na + nb) * (na + nb - 1.0);
    }
    for (int h = 0; h < H; h++) r += 2L * GAA(h) * GAA(h);
    for (int h = 0; h < H; h++) r += (long)GAB(h) * GAB(h);
    for (int spin = 0; spin < 2; spin++) /* D1 + Q1 = I, alpha then beta */
        for (int h = 0; h < H; h++) {
            int m = p->amopi[h];
            for (int i = 0; i < m; i++) b[r + i * m + i] = 1.0;
            r += (long)m * m;
        }
    for (int s = 0; s < (p->constrain_sz ? 4 : 2); s++)
        for (int h = 0; h < H; h++) r += (long)p->amopi[h] * p->amopi[h];
    if (p->constrain_spin) {
        double ms = (p->multiplicity - 1.0) / 2.0;
        b[r++] = 0.5 * (na + nb) + ms * ms - ms * (ms + 1.0);
        if (p->nalpha == p->nbeta) {
            for (int h = 0; h < H; h++) r += (long)p->amopi[h] * p->amopi[h];
            for (int h = 0; h < H; h++) r += 3L * GAA(h) * GAA(h);
            for (int h = 0; h < H; h++) r += 2L * GAB(h) * GAB(h);
        } else {
            for (int h = 0; h < H; h++) r += 4L * GAB(h) * GAB(h);
        }
        if (p->constrain_g2) r += 2L * GAB(0);
    }
    if (p->constrain_q2) {
        for (int h = 0; h < H; h++) {
            for (int ij = 0; ij < GAB(h); ij++)
                for (int kl = 0; kl < GAB(h); kl++)
                    b[r + (long)ij * GAB(h) + kl] =
                        -(double)((BAB0(h, ij) == BAB0(h, kl)) && (BAB1(h, ij) == BAB1(h, kl)));
            r += (long)GAB(h) * GAB(h);
        }
        for (int spin = 0; spin < 2; spin++)
            for (int h = 0; h < H; h++) {
                for (int ij = 0; ij < GAA(h); ij++) {
                    int i = BAA0(h, ij), j = BAA1(h, ij);
                    for (int kl = 0; kl < GAA(h); kl++) {
                        int k = BAA0(h, kl), l = BAA1(h, kl);
                        b[r + (long)ij * GAA(h) + kl] = -(double)(i == k && j == l) + (double)(i == l && j == k);
                    }
                }
                r += (long)GAA(h) * GAA(h);
            }
    }
    /* G2 rows have b = 0 */
}

/* ---- c ------------------------------------------------------------------ */
static inline long g4(int n, int a, int b, int c, int d) {
    return (((long)a * n + b) * n + c) * n + d;
}

/* c[D1a] = c[D1b] = h (tei.cc:245-271 with no core orbitals);
 * c[D2ab](ij,kl) = (ik|jl); c[D2aa/bb](ij,kl) = (ik|jl) - (il|jk)
 * (v2rdm_solver.cc:164-193, tei.cc:139-199). */
void orc_pack_c(const orc_problem *p, const double *hm, const double *g, double *c) {
    int n = NMO, H = p->nirrep;
    memset(c, 0, sizeof(double) * p->n_primal);
    for (int h = 0; h < H; h++) {
        int m = p->amopi[h], o = p->pitzer_offset[h];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                double v = hm[(long)(i + o) * n + (j + o)];
                c[p->d1aoff[h] + (long)i * m + j] = v;
                c[p->d1boff[h] + (long)i * m + j] = v;
            }
        for (int ij = 0; ij < GAB(h); ij++) {
            int i = BAB0(h, ij), j = BAB1(h, ij);
            for (int kl = 0; kl < GAB(h); kl++) {
                int k = BAB0(h, kl), l = BAB1(h, kl);
                c[p->d2aboff[h] + (long)ij * GAB(h) + kl] = g[g4(n, i, k, j, l)];
            }
        }
        for (int ij = 0; ij < GAA(h); ij++) {
            int i = BAA0(h, ij), j = BAA1(h, ij);
            for (int kl = 0; kl < GAA(h); kl++) {
                int k = BAA0(h, kl), l = BAA1(h, kl);
                double v = g[g4(n, i, k, j, l)] - g[g4(n, i, l, j, k)];
                c[p->d2aaoff[h] + (long)ij * GAA(h) + kl] = v;
                c[p->d2bboff[h] + (long)ij * GAA(h) + kl] = v;
            }
        }
    }
}

/* Exact transpose of orc_pack_c. */
void orc_pack_c_adjoint(const orc_problem *p, const double *xb, double *hb, double *gb) {
    int n = NMO, H = p->nirrep;
    memset(hb, 0, sizeof(double) * n * n);
    memset(gb, 0, sizeof(double) * (long)n * n * n * n);
    for (int h = 0; h < H; h++) {
        int m = p->amopi[h], o = p->pitzer_offset[h];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++)
                hb[(long)(i + o) * n + (j + o)] += xb[p->d1aoff[h] + (long)i * m + j]
                                                 + xb[p->d1boff[h] + (long)i * m + j];
        for (int ij = 0; ij < GAB(h); ij++) {
            int i = BAB0(h, ij), j = BAB1(h, ij);
            for (int kl = 0; kl < GAB(h); kl++) {
                int k = BAB0(h, kl), l = BAB1(h, kl);
                gb[g4(n, i, k, j, l)] += xb[p->d2aboff[h] + (long)ij * GAB(h) + kl];
            }
        }
        for (int ij = 0; ij < GAA(h); ij++) {
            int i = BAA0(h, ij), j = BAA1(h, ij);
            for (int kl = 0; kl < GAA(h); kl++) {
                int k = BAA0(h, kl), l = BAA1(h, kl);
                double v = xb[p->d2aaoff[h] + (long)ij * GAA(h) + kl] + xb[p->d2bboff[h] + (long)ij * GAA(h) + kl];
                gb[g4(n, i, k, j, l)] += v;
                gb[g4(n, i, l, j, k)] -= v;
            }
        }
    }
}

/* oei.cc:76-95: -t between ring neighbours; tei.cc:285-291: U on site. */
void orc_hubbard_integrals(int n, double t, double U, double *hm, double *g) {
    memset(hm, 0, sizeof(double) * n * n);
    memset(g, 0, sizeof(double) * (long)n * n * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int d = abs(i - j);
            if (d == 1 || d == n - 1) hm[i * n + j] = -t;
        }
    for (int i = 0; i < n; i++) g[g4(n, i, i, i, i)] = U;
}

/* psi4's INDEX: packed lower triangle (the same formula as INDEX2,
 * misc/integraltransform_sort_so_tpdm.cc:41). */
static long tri_index(long i, long j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

/* tei.cc:280-283: C_DDOT over Q with stride ntri. */
static double df_tei(const double *qmo, int nq, long ntri, int i, int j, int k, int l) {
    const double *a = qmo + tri_index(i, j), *b = qmo + tri_index(k, l);
    double s = 0.0;
    for (int Q = 0; Q < nq; Q++) s += a[(long)Q * ntri] * b[(long)Q * ntri];
    return s;
}

/* oei_full_sym_ keeps i <= j (tei.cc:124-132) and is read through INDEX. */
static double oei_up(const double *h, int n, int i, int j) {
    return i <= j ? h[(long)i * n + j] : h[(long)j * n + i];
}

void orc_df_integrals(int nfull, const double *h_full, int nq, const double *qmo, int ncore, const int *core,
                      int nact, const int *active, double *h_act, double *g_act, double *efzc) {
    const long ntri = (long)nfull * (nfull + 1) / 2;
    /* FrozenCoreEnergy, tei.cc:204-227 */
    double e = 0.0;
    for (int a = 0; a < ncore; a++) {
        const int i = core[a];
        e += 2.0 * oei_up(h_full, nfull, i, i);
        for (int b = 0; b < ncore; b++) {
            const int j = core[b];
            const double dum1 = df_tei(qmo, nq, ntri, i, i, j, j);
            const double dum2 = df_tei(qmo, nq, ntri, i, j, i, j);
            e += 2.0 * dum1 - dum2;
        }
    }
    *efzc = e;
    /* core repulsion in the one-electron integrals, tei.cc:229-268 */
    for (int p = 0; p < nact; p++)
        for (int q = 0; q < nact; q++) {
            const int i = active[p], j = active[q];
            double dum = 0.0;
            for (int c = 0; c < ncore; c++) {
                const int k = core[c];
                const double dum1 = df_tei(qmo, nq, ntri, i, j, k, k);
                const double dum2 = df_tei(qmo, nq, ntri, i, k, j, k);
                dum += 2.0 * dum1 - dum2;
            }
            h_act[(long)p * nact + q] = oei_up(h_full, nfull, i, j) + dum;
        }
    /* the (pq|rs) that RepackIntegrals packs, tei.cc:155-197 */
    const long n = nact;
    for (int p = 0; p < nact; p++)
        for (int q = 0; q < nact; q++)
            for (int r = 0; r < nact; r++)
                for (int s = 0; s < nact; s++)
                    g_act[((p * n + q) * n + r) * n + s] =
                        df_tei(qmo, nq, ntri, active[p], active[q], active[r], active[s]);
}

/* ---- FCIDUMP records (fcidump.cc:60-243) ------------------------------------ */

void orc_fcidump_map(int nfull, int nirrep, const int *full_irreps, const int *space, int *map) {
    int count = 0;
    for (int sp = 0; sp < 3; sp++)
        for (int h = 0; h < nirrep; h++)
            for (int p = 0; p < nfull; p++)
                if (space[p] == sp && (full_irreps ? full_irreps[p] : 0) == h) map[p] = ++count;
}

static void put_rec(double *val, int *lab, long k, double v, int a, int b, int c, int d) {
    if (!val) return;
    val[k] = v;
    lab[4 * k] = a;
    lab[4 * k + 1] = b;
    lab[4 * k + 2] = c;
    lab[4 * k + 3] = d;
}

long orc_fcidump_integrals(int nfull, const int *full_irreps, const int *map, const double *h_full, int nq,
                           const double *qmo, double enuc, double *val, int *lab) {
    const long ntri = (long)nfull * (nfull + 1) / 2;
#define FIRR(p) (full_irreps ? full_irreps[(p)] : 0)
    long k = 0;
    for (int p = 0; p < nfull; p++)
        for (int q = p; q < nfull; q++) {
            const int hpq = FIRR(p) ^ FIRR(q);
            const long pq = tri_index(p, q);
            for (int r = 0; r < nfull; r++)
                for (int s = r; s < nfull; s++) {
                    if ((FIRR(r) ^ FIRR(s)) != hpq) continue;
                    if (pq > tri_index(r, s)) continue;
                    if (val) put_rec(val, lab, k, df_tei(qmo, nq, ntri, p, q, r, s), map[p], map[q], map[r], map[s]);
                    k++;
                }
        }
    int nirrep = 1;
    for (int p = 0; p < nfull; p++)
        while (FIRR(p) >= nirrep) nirrep *= 2;
    for (int h = 0; h < nirrep; h++)
        for (int p = 0; p < nfull; p++) {
            if (FIRR(p) != h) continue;
            for (int q = 0; q < nfull; q++) {
                if (FIRR(q) != h) continue;
                put_rec(val, lab, k, h_full[(long)p * nfull + q], map[p], map[q], 0, 0);
                k++;
            }
        }
#undef FIRR
    put_rec(val, lab, k, enuc, 0, 0, 0, 0);
    return k + 1;
}

long orc_fcidump_rdms(const orc_problem *p, const double *x, const int *act_full, const int *map, int ninact,
                      double *val, int *lab) {
    long k = 0;
    for (int h = 0; h < p->nirrep; h++)
        for (int ij = 0; ij < GAB(h); ij++) {
            const int i = BAB0(h, ij), j = BAB1(h, ij), ji = IAB(h, j, i);
            for (int kl = 0; kl < GAB(h); kl++) {
                const int kk = BAB0(h, kl), l = BAB1(h, kl), lk = IAB(h, l, kk);
                double dum = x[p->d2aboff[h] + (long)ij * GAB(h) + kl] + x[p->d2aboff[h] + (long)ji * GAB(h) + lk];
                if (i != j && kk != l) {
                    const int ija = IAA(h, i, j), kla = IAA(h, kk, l);
                    int sg = 1;
                    if (i > j) sg = -sg;
                    if (kk > l) sg = -sg;
                    dum += sg * x[p->d2aaoff[h] + (long)ija * GAA(h) + kla];
                    dum += sg * x[p->d2bboff[h] + (long)ija * GAA(h) + kla];
                }
                put_rec(val, lab, k, dum, map[act_full[i]] - ninact, map[act_full[j]] - ninact,
                        map[act_full[kk]] - ninact, map[act_full[l]] - ninact);
                k++;
            }
        }
    for (int h = 0; h < p->nirrep; h++)
        for (int ij = 0; ij < GAB(h); ij++) {
            const int i = BAB0(h, ij), j = BAB1(h, ij);
            const int hi = SYM(i), hj = SYM(j);
            if (hi != hj) continue;
            const int ii = i - p->pitzer_offset[hi], jj = j - p->pitzer_offset[hj];
            const double dum = x[p->d1aoff[hi] + (long)ii * p->amopi[hi] + jj] + x[p->d1boff[hi] + (long)ii * p->amopi[hi] + jj];
            put_rec(val, lab, k, dum, map[act_full[i]] - ninact, map[act_full[j]] - ninact, 0, 0);
            k++;
        }
    return k;
}

/* ---- T1 rows (t1.cc:419-793) ------------------------------------------------- */
/* Row order T1aab, T1bba, T1aaa, T1bbb (the order T1_constraints_Au writes
 * them with its running offset); every row is - T1(ijk,lmn) plus the Q2, D2,
 * G2 and D1 elements t1.cc names, with the same Kronecker conditions.  The
 * terms are handed to `term` in the reference's order, so that Au sums them
 * in that order and A^T u scatters the same terms. */
typedef void (*t1_term_fn)(long row, long col, double coef, void *ctx);

static long t1_rows(const orc_problem *p, long r, t1_term_fn term, void *ctx) {
    const int H = p->nirrep;
#define T3(arr, h, a, c) (p->arr[(h)][3 * (a) + (c)])
#define D1(off, a, b) ((off)[SYM(a)] + (long)((a) - p->pitzer_offset[SYM(a)]) * p->amopi[SYM(a)] + ((b) - p->pitzer_offset[SYM(a)]))
    for (int s = 0; s < 2; s++) /* 0: T1aab, 1: T1bba */
        for (int h = 0; h < H; h++) {
            const int t = p->trip_aab[h];
            const long toff = s == 0 ? p->t1aaboff[h] : p->t1bbaoff[h];
            for (int ijk = 0; ijk < t; ijk++) {
                const int i = T3(bas_aab, h, ijk, 0), j = T3(bas_aab, h, ijk, 1), k = T3(bas_aab, h, ijk, 2);
                for (int lmn = 0; lmn < t; lmn++) {
                    const int l = T3(bas_aab, h, lmn, 0), m = T3(bas_aab, h, lmn, 1), n = T3(bas_aab, h, lmn, 2);
                    const long row = r + (long)ijk * t + lmn;
                    term(row, toff + (long)ijk * t + lmn, -1.0, ctx);
                    if (k == n) {
                        const int hij = SYM(i) ^ SYM(j);
                        const long q2 = s == 0 ? p->q2aaoff[hij] : p->q2bboff[hij];
                        term(row, q2 + (long)IAA(hij, i, j) * GAA(hij) + IAA(hij, l, m), 1.0, ctx);
                    }
                    if (j == l) {
                        const int hki = SYM(k) ^ SYM(i);
                        const long nm = s == 0 ? IAB(hki, m, n) : IAB(hki, n, m);
                        const long ki = s == 0 ? IAB(hki, i, k) : IAB(hki, k, i);
                        term(row, p->d2aboff[hki] + nm * GAB(hki) + ki, -1.0, ctx);
                    }
                    if (l == i) {
                        const int hkj = SYM(k) ^ SYM(j);
                        const long nm = s == 0 ? IAB(hkj, m, n) : IAB(hkj, n, m);
                        const long kj = s == 0 ? IAB(hkj, j, k) : IAB(hkj, k, j);
                        term(row, p->d2aboff[hkj] + nm * GAB(hkj) + kj, 1.0, ctx);
                    }
                    if (j == m) {
                        const int hni = SYM(n) ^ SYM(i);
                        const long g2 = s == 0 ? p->g2baoff[hni] : p->g2aboff[hni];
                        term(row, g2 + (long)IAB(hni, n, i) * GAB(hni) + IAB(hni, k, l), -1.0, ctx);
                    }
                    if (i == m) {
                        const int hkl = SYM(k) ^ SYM(l);
                        const long g2 = s == 0 ? p->g2baoff[hkl] : p->g2aboff[hkl];
                        term(row, g2 + (long)IAB(hkl, n, j) * GAB(hkl) + IAB(hkl, k, l), 1.0, ctx);
                    }
                }
            }
            r += (long)t * t;
        }
    for (int s = 0; s < 2; s++) /* 0: T1aaa, 1: T1bbb */
        for (int h = 0; h < H; h++) {
            const int t = p->trip_aaa[h];
            const long toff = s == 0 ? p->t1aaaoff[h] : p->t1bbboff[h];
            const long *q2o = s == 0 ? p->q2aaoff : p->q2bboff;
            const long *d2o = s == 0 ? p->d2aaoff : p->d2bboff;
            const long *d1o = s == 0 ? p->d1aoff : p->d1boff;
            for (int ijk = 0; ijk < t; ijk++) {
                const int i = T3(bas_aaa, h, ijk, 0), j = T3(bas_aaa, h, ijk, 1), k = T3(bas_aaa, h, ijk, 2);
                for (int lmn = 0; lmn < t; lmn++) {
                    const int l = T3(bas_aaa, h, lmn, 0), m = T3(bas_aaa, h, lmn, 1), n = T3(bas_aaa, h, lmn, 2);
                    const long row = r + (long)ijk * t + lmn;
                    /* G2aa element (a, b) of the aa quadrant (s = 0) or bb quadrant (s = 1) */
#define G2Q(hh, a, b) (p->g2aaoff[(hh)] + (long)((a) + s * GAB(hh)) * 2 * GAB(hh) + ((b) + s * GAB(hh)))
                    term(row, toff + (long)ijk * t + lmn, -1.0, ctx);
                    if (k == n) {
                        const int hij = SYM(i) ^ SYM(j);
                        term(row, q2o[hij] + (long)IAA(hij, i, j) * GAA(hij) + IAA(hij, l, m), 1.0, ctx);
                    }
                    if (j == n) {
                        const int hik = SYM(i) ^ SYM(k), hlm = SYM(l) ^ SYM(m);
                        if (hik == hlm) term(row, q2o[hik] + (long)IAA(hik, i, k) * GAA(hik) + IAA(hik, l, m), -1.0, ctx);
                    }
                    if (i == n) {
                        const int hjk = SYM(j) ^ SYM(k), hlm = SYM(l) ^ SYM(m);
                        if (hjk == hlm) term(row, q2o[hjk] + (long)IAA(hjk, j, k) * GAA(hjk) + IAA(hjk, l, m), 1.0, ctx);
                    }
                    if (l == k) {
                        const int hnm = SYM(n) ^ SYM(m), hji = SYM(j) ^ SYM(i);
                        if (hji == hnm) term(row, d2o[hji] + (long)IAA(hji, n, m) * GAA(hji) + IAA(hji, j, i), 1.0, ctx);
                    }
                    if (j == l) {
                        const int hki = SYM(k) ^ SYM(i);
                        term(row, d2o[hki] + (long)IAA(hki, n, m) * GAA(hki) + IAA(hki, k, i), -1.0, ctx);
                    }
                    if (l == i) {
                        const int hkj = SYM(k) ^ SYM(j);
                        term(row, d2o[hkj] + (long)IAA(hkj, n, m) * GAA(hkj) + IAA(hkj, k, j), 1.0, ctx);
                    }
                    if (k == m) {
                        if (j == l) term(row, D1(d1o, n, i), -1.0, ctx);
                        const int hni = SYM(n) ^ SYM(i);
                        term(row, G2Q(hni, IAB(hni, n, i), IAB(hni, j, l)), 1.0, ctx);
                    }
                    if (j == m) {
                        if (k == l) term(row, D1(d1o, n, i), 1.0, ctx);
                        const int hni = SYM(n) ^ SYM(i);
                        term(row, G2Q(hni, IAB(hni, n, i), IAB(hni, k, l)), -1.0, ctx);
                    }
                    if (i == m) {
                        if (k == l) term(row, D1(d1o, n, j), -1.0, ctx);
                        const int hkl = SYM(k) ^ SYM(l);
                        term(row, G2Q(hkl, IAB(hkl, n, j), IAB(hkl, k, l)), 1.0, ctx);
                    }
#undef G2Q
                }
            }
            r += (long)t * t;
        }
#undef T3
#undef D1
    return r;
}

/* ---- T2 rows (t2.cc:1236-1843, T2_constraints_Au_slow) ------------------ */
/* Row order T2aab, T2bba, T2aaa (one block of dimension trip_aab + trip_aba:
 * aab x aab, aab x aba, aba x aab, aba x aba), T2bbb (the same four parts);
 * every row is - T2(ijk,lmn) plus the D2 and D1 elements t2.cc names. */
static long t2_rows(const orc_problem *p, long r, t1_term_fn term, void *ctx) {
    const int H = p->nirrep;
#define T3(arr, h, a, c) (p->arr[(h)][3 * (a) + (c)])
#define D1(off, a, b) ((off)[SYM(a)] + (long)((a) - p->pitzer_offset[SYM(a)]) * p->amopi[SYM(a)] + ((b) - p->pitzer_offset[SYM(a)]))
#define DAB(hh, a, b) (p->d2aboff[(hh)] + (long)(a) * GAB(hh) + (b))
#define DSS(offs, hh, a, b) ((offs)[(hh)] + (long)(a) * GAA(hh) + (b))
    /* signed D2aa / D2bb term of the T2aaa-type parts: s (x > y) (u > v) on
     * the pair (x, y), (u, v), skipped when a pair repeats an orbital */
#define SIGNED(offs, x, y, u, v, c0)                                                          \
    do {                                                                                      \
        if ((x) != (y) && (u) != (v)) {                                                       \
            const int hh = SYM(x) ^ SYM(y);                                                   \
            double sg = (c0);                                                                 \
            if ((x) > (y)) sg = -sg;                                                          \
            if ((u) > (v)) sg = -sg;                                                          \
            term(row, DSS(offs, hh, IAA(hh, x, y), IAA(hh, u, v)), sg, ctx);                  \
        }                                                                                     \
    } while (0)
    for (int s = 0; s < 2; s++) /* 0: T2aab, 1: T2bba */
        for (int h = 0; h < H; h++) {
            const int t = p->trip_aab[h];
            const long toff = s == 0 ? p->t2aaboff[h] : p->t2bbaoff[h];
            const long *d2s = s == 0 ? p->d2aaoff : p->d2bboff;
            const long *d1o = s == 0 ? p->d1boff : p->d1aoff;
            for (int ijk = 0; ijk < t; ijk++) {
                const int i = T3(bas_aab, h, ijk, 0), j = T3(bas_aab, h, ijk, 1), k = T3(bas_aab, h, ijk, 2);
                for (int lmn = 0; lmn < t; lmn++) {
                    const int l = T3(bas_aab, h, lmn, 0), m = T3(bas_aab, h, lmn, 1), n = T3(bas_aab, h, lmn, 2);
                    const long row = r + (long)ijk * t + lmn;
                    term(row, toff + (long)ijk * t + lmn, -1.0, ctx);
                    if (k == n) {
                        const int hij = SYM(i) ^ SYM(j);
                        term(row, DSS(d2s, hij, IAA(hij, i, j), IAA(hij, l, m)), 1.0, ctx);
                    }
                    if (j == m && i == l) term(row, D1(d1o, n, k), 1.0, ctx);
                    if (j == l && i == m) term(row, D1(d1o, n, k), -1.0, ctx);
                    if (i == l) {
                        const int hh = SYM(n) ^ SYM(j);
                        term(row, s == 0 ? DAB(hh, IAB(hh, j, n), IAB(hh, m, k)) : DAB(hh, IAB(hh, n, j), IAB(hh, k, m)), -1.0, ctx);
                    }
                    if (j == l) {
                        const int hh = SYM(n) ^ SYM(i);
                        term(row, s == 0 ? DAB(hh, IAB(hh, i, n), IAB(hh, m, k)) : DAB(hh, IAB(hh, n, i), IAB(hh, k, m)), 1.0, ctx);
                    }
                    if (i == m) {
                        const int hh = SYM(n) ^ SYM(j);
                        term(row, s == 0 ? DAB(hh, IAB(hh, j, n), IAB(hh, l, k)) : DAB(hh, IAB(hh, n, j), IAB(hh, k, l)), 1.0, ctx);
                    }
                    if (j == m) {
                        const int hh = SYM(n) ^ SYM(i);
                        term(row, s == 0 ? DAB(hh, IAB(hh, i, n), IAB(hh, l, k)) : DAB(hh, IAB(hh, n, i), IAB(hh, k, l)), -1.0, ctx);
                    }
                }
            }
            r += (long)t * t;
        }
    for (int s = 0; s < 2; s++) /* 0: T2aaa (+ abb), 1: T2bbb (+ baa) */
        for (int h = 0; h < H; h++) {
            const int ta = p->trip_aab[h], tb = p->trip_aba[h], D = ta + tb;
            const long toff = s == 0 ? p->t2aaaoff[h] : p->t2bbboff[h];
            const long *d2s = s == 0 ? p->d2aaoff : p->d2bboff;  /* same-spin D2 of the aab part */
            const long *d2o = s == 0 ? p->d2bboff : p->d2aaoff;  /* same-spin D2 of the aba part */
            const long *d1s = s == 0 ? p->d1aoff : p->d1boff;
            const long *d1o = s == 0 ? p->d1boff : p->d1aoff;
            /* aab x aab */
            for (int ijk = 0; ijk < ta; ijk++) {
                const int i = T3(bas_aab, h, ijk, 0), j = T3(bas_aab, h, ijk, 1), k = T3(bas_aab, h, ijk, 2);
                for (int lmn = 0; lmn < ta; lmn++) {
                    const int l = T3(bas_aab, h, lmn, 0), m = T3(bas_aab, h, lmn, 1), n = T3(bas_aab, h, lmn, 2);
                    const long id = (long)ijk * D + lmn, row = r + id;
                    term(row, toff + id, -1.0, ctx);
                    if (k == n) {
                        const int hij = SYM(i) ^ SYM(j);
                        term(row, DSS(d2s, hij, IAA(hij, i, j), IAA(hij, l, m)), 1.0, ctx);
                    }
                    if (j == m && i == l) term(row, D1(d1s, n, k), 1.0, ctx);
                    if (j == l && i == m) term(row, D1(d1s, n, k), -1.0, ctx);
                    if (i == l) SIGNED(d2s, n, j, k, m, -1.0);
                    if (j == l) SIGNED(d2s, n, i, k, m, 1.0);
                    if (i == m) SIGNED(d2s, n, j, k, l, 1.0);
                    if (j == m) SIGNED(d2s, n, i, k, l, -1.0);
                }
            }
            /* aab x aba */
            for (int ijk = 0; ijk < ta; ijk++) {
                const int i = T3(bas_aab, h, ijk, 0), j = T3(bas_aab, h, ijk, 1), k = T3(bas_aab, h, ijk, 2);
                for (int lmn = 0; lmn < tb; lmn++) {
                    const int l = T3(bas_aba, h, lmn, 0), m = T3(bas_aba, h, lmn, 1), n = T3(bas_aba, h, lmn, 2);
                    const long id = (long)ijk * D + (lmn + ta), row = r + id;
                    term(row, toff + id, -1.0, ctx);
                    if (i == l) {
                        const int hh = SYM(j) ^ SYM(n);
                        term(row, s == 0 ? DAB(hh, IAB(hh, j, n), IAB(hh, k, m)) : DAB(hh, IAB(hh, n, j), IAB(hh, m, k)), 1.0, ctx);
                    }
                    if (j == l) {
                        const int hh = SYM(i) ^ SYM(n);
                        term(row, s == 0 ? DAB(hh, IAB(hh, i, n), IAB(hh, k, m)) : DAB(hh, IAB(hh, n, i), IAB(hh, m, k)), -1.0, ctx);
                    }
                }
            }
            /* aba x aab */
            for (int ijk = 0; ijk < tb; ijk++) {
                const int i = T3(bas_aba, h, ijk, 0), j = T3(bas_aba, h, ijk, 1), k = T3(bas_aba, h, ijk, 2);
                for (int lmn = 0; lmn < ta; lmn++) {
                    const int l = T3(bas_aab, h, lmn, 0), m = T3(bas_aab, h, lmn, 1), n = T3(bas_aab, h, lmn, 2);
                    const long id = (long)(ijk + ta) * D + lmn, row = r + id;
                    term(row, toff + id, -1.0, ctx);
                    if (i == l) {
                        const int hh = SYM(j) ^ SYM(n);
                        term(row, s == 0 ? DAB(hh, IAB(hh, n, j), IAB(hh, m, k)) : DAB(hh, IAB(hh, j, n), IAB(hh, k, m)), 1.0, ctx);
                    }
                    if (i == m) {
                        const int hh = SYM(j) ^ SYM(n);
                        term(row, s == 0 ? DAB(hh, IAB(hh, n, j), IAB(hh, l, k)) : DAB(hh, IAB(hh, j, n), IAB(hh, k, l)), -1.0, ctx);
                    }
                }
            }
            /* aba x aba */
            for (int ijk = 0; ijk < tb; ijk++) {
                const int i = T3(bas_aba, h, ijk, 0), j = T3(bas_aba, h, ijk, 1), k = T3(bas_aba, h, ijk, 2);
                for (int lmn = 0; lmn < tb; lmn++) {
                    const int l = T3(bas_aba, h, lmn, 0), m = T3(bas_aba, h, lmn, 1), n = T3(bas_aba, h, lmn, 2);
                    const long id = (long)(ijk + ta) * D + (lmn + ta), row = r + id;
                    term(row, toff + id, -1.0, ctx);
                    if (k == n) {
                        const int hij = SYM(i) ^ SYM(j);
                        term(row, s == 0 ? DAB(hij, IAB(hij, i, j), IAB(hij, l, m)) : DAB(hij, IAB(hij, j, i), IAB(hij, m, l)), 1.0, ctx);
                    }
                    if (j == m && i == l) term(row, D1(d1o, n, k), 1.0, ctx);
                    if (i == l) SIGNED(d2o, n, j, k, m, -1.0);
                    if (j == m) {
                        const int hh = SYM(n) ^ SYM(i);
                        term(row, s == 0 ? DAB(hh, IAB(hh, i, n), IAB(hh, l, k)) : DAB(hh, IAB(hh, n, i), IAB(hh, k, l)), -1.0, ctx);
                    }
                }
            }
            r += (long)D * D;
        }
#undef SIGNED
#undef DSS
#undef DAB
#undef D1
#undef T3
    return r;
}

/* ---- D3 rows (d3.cc:45-406) ------------------------------------------------- */
/* With CONSTRAIN_SZ: D3aaa -> D2aa, D3bbb -> D2bb, D3aab -> D2aa, D3bba ->
 * D2bb, D3aab -> D2ab, D3bba -> D2ab; without: (D3aaa + D3aab) -> D2aa,
 * (D3bbb + D3bba) -> D2bb, (D3aab + D3bba) -> D2ab; then for singlets with
 * CONSTRAIN_SPIN: D3aab = D3bba and D3aaa, D3bbb from D3aab, D3bba. */
static long d3_rows(const orc_problem *p, long r, t1_term_fn term, void *ctx) {
    const int H = p->nirrep, n = NMO;
    const double na = p->nalpha, nb = p->nbeta;
#define IAAA(hh, a, b, c) (p->ibas_aaa[(hh)][((long)(a) * n + (b)) * n + (c)])
#define IAAB(hh, a, b, c) (p->ibas_aab[(hh)][((long)(a) * n + (b)) * n + (c)])
    /* sum over p of the D3 same-spin (aaa) element (ijp, klp), sign of
     * d3.cc's (i < p < j), (k < p < l) rule */
#define AAA_TERM(off3, c)                                                                     \
    do {                                                                                      \
        if (i != q && j != q && k != q && l != q) {                                           \
            const int h2 = h ^ SYM(q);                                                        \
            double sg = (c);                                                                  \
            if (i < q && q < j) sg = -sg;                                                     \
            if (k < q && q < l) sg = -sg;                                                     \
            term(row, (off3)[h2] + (long)IAAA(h2, i, j, q) * p->trip_aaa[h2] + IAAA(h2, k, l, q), sg, ctx); \
        }                                                                                     \
    } while (0)
    const int nsame = p->constrain_sz ? 2 : 1;
    /* same-spin D2 rows: sz: aaa->aa, bbb->bb, aab->aa, bba->bb; else (aaa+aab)->aa, (bbb+bba)->bb */
    for (int pass = 0; pass < 2 * nsame; pass++) {
        const int spin = pass % 2;        /* 0: alpha rows (D2aa), 1: beta rows (D2bb) */
        const int part = p->constrain_sz ? pass / 2 : -1;  /* 0: from D3aaa/bbb, 1: from D3aab/bba, -1: both */
        const long *d2o = spin ? p->d2bboff : p->d2aaoff;
        const long *o3s = spin ? p->d3bbboff : p->d3aaaoff;
        const long *o3m = spin ? p->d3bbaoff : p->d3aaboff;
        double coef;
        if (part == 0) coef = (spin ? nb : na) - 2.0;
        else if (part == 1) coef = spin ? na : nb;
        else coef = na + nb - 2.0;
        for (int h = 0; h < H; h++) {
            const int g = GAA(h);
            for (int ij = 0; ij < g; ij++) {
                const int i = BAA0(h, ij), j = BAA1(h, ij);
                for (int kl = 0; kl < g; kl++) {
                    const int k = BAA0(h, kl), l = BAA1(h, kl);
                    const long row = r + (long)ij * g + kl;
                    term(row, d2o[h] + (long)ij * g + kl, coef, ctx);
                    for (int q = 0; q < n; q++) {
                        const int h2 = h ^ SYM(q);
                        if (part != 0)
                            term(row, o3m[h2] + (long)IAAB(h2, i, j, q) * p->trip_aab[h2] + IAAB(h2, k, l, q), -1.0, ctx);
                        if (part != 1) AAA_TERM(o3s, -1.0);
                    }
                }
            }
            r += (long)g * g;
        }
    }
    /* D2ab rows: sz: from D3aab, then from D3bba; else one row from both */
    for (int pass = 0; pass < nsame; pass++) {
        double coef = p->constrain_sz ? (pass == 0 ? na - 1.0 : nb - 1.0) : na + nb - 2.0;
        for (int h = 0; h < H; h++) {
            const int g = GAB(h);
            for (int ij = 0; ij < g; ij++) {
                const int i = BAB0(h, ij), j = BAB1(h, ij);
                for (int kl = 0; kl < g; kl++) {
                    const int k = BAB0(h, kl), l = BAB1(h, kl);
                    const long row = r + (long)ij * g + kl;
                    term(row, p->d2aboff[h] + (long)ij * g + kl, coef, ctx);
                    for (int q = 0; q < n; q++) {
                        const int h2 = h ^ SYM(q);
                        if ((!p->constrain_sz || pass == 0) && i != q && k != q) {
                            double sg = -1.0;
                            if (q < i) sg = -sg;
                            if (q < k) sg = -sg;
                            term(row, p->d3aaboff[h2] + (long)IAAB(h2, i, q, j) * p->trip_aab[h2] + IAAB(h2, k, q, l), sg, ctx);
                        }
                        if ((!p->constrain_sz || pass == 1) && j != q && l != q) {
                            double sg = -1.0;
                            if (q < j) sg = -sg;
                            if (q < l) sg = -sg;
                            term(row, p->d3bbaoff[h2] + (long)IAAB(h2, j, q, i) * p->trip_aab[h2] + IAAB(h2, l, q, k), sg, ctx);
                        }
                    }
                }
            }
            r += (long)g * g;
        }
    }
    if (p->constrain_spin && p->nalpha == p->nbeta) {
        for (int h = 0; h < H; h++) { /* D3aab - D3bba = 0 */
            const long t2 = (long)p->trip_aab[h] * p->trip_aab[h];
            for (long e = 0; e < t2; e++) {
                term(r + e, p->d3aaboff[h] + e, 1.0, ctx);
                term(r + e, p->d3bbaoff[h] + e, -1.0, ctx);
            }
            r += t2;
        }
        for (int s = 0; s < 2; s++) /* D3aaa from D3aab, D3bbb from D3bba */
            for (int h = 0; h < H; h++) {
                const int t = p->trip_aaa[h], ta = p->trip_aab[h];
                const long o3 = s ? p->d3bbboff[h] : p->d3aaaoff[h], om = s ? p->d3bbaoff[h] : p->d3aaboff[h];
                for (int pqr = 0; pqr < t; pqr++) {
                    const int a = p->bas_aaa[h][3 * pqr], b = p->bas_aaa[h][3 * pqr + 1], c = p->bas_aaa[h][3 * pqr + 2];
                    const long rows3[3] = {IAAB(h, a, b, c), IAAB(h, a, c, b), IAAB(h, b, c, a)};
                    for (int stv = 0; stv < t; stv++) {
                        const int s1 = p->bas_aaa[h][3 * stv], t1 = p->bas_aaa[h][3 * stv + 1], v1 = p->bas_aaa[h][3 * stv + 2];
                        const long cols3[3] = {IAAB(h, s1, t1, v1), IAAB(h, s1, v1, t1), IAAB(h, t1, v1, s1)};
                        const long row = r + (long)pqr * t + stv;
                        term(row, o3 + (long)pqr * t + stv, 1.0, ctx);
                        for (int x = 0; x < 3; x++)
                            for (int y = 0; y < 3; y++) {
                                /* signs - + - / + - + / - + - of d3.cc:360-371 */
                                const double sg = ((x + y) % 2 == 0 ? -1.0 : 1.0) / 3.0;
                                term(row, om + rows3[x] * ta + cols3[y], sg, ctx);
                            }
                    }
                }
                r += (long)t * t;
            }
    }
#undef AAA_TERM
#undef IAAB
#undef IAAA
    return r;
}

struct t1_ctx { const double *in; double *out; };
static void t1_gather(long row, long col, double coef, void *c) {
    struct t1_ctx *x = (struct t1_ctx *)c;
    x->out[row] += coef * x->in[col];
}
static void t1_scatter(long row, long col, double coef, void *c) {
    struct t1_ctx *x = (struct t1_ctx *)c;
    x->out[col] += coef * x->in[row];
}

/* ---- A u ---------------------------------------------------------------- */
/* v2rdm_solver.cc:2995-3050 (dispatch order D2, spin, Q2, G2, then T1 at
 * :3015-3016). */
void orc_Au(const orc_problem *p, const double *u, double *A) {
    const int n = NMO, H = p->nirrep;
    const double na = p->nalpha, nb = p->nbeta;
    memset(A, 0, sizeof(double) * p->n_dual);
    long r = 0;

    /* d2.cc:316-380 traces */
    if (p->constrain_sz) {
        double tab = 0.0, taa = 0.0, tbb = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int h = SYM(i) ^ SYM(j);
                int ij = IAB(h, i, j);
                tab += u[p->d2aboff[h] + (long)ij * GAB(h) + ij];
            }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                if (i == j) continue;
                int h = SYM(i) ^ SYM(j);
                int ij = IAA(h, i, j);
                taa += u[p->d2aaoff[h] + (long)ij * GAA(h) + ij];
                tbb += u[p->d2bboff[h] + (long)ij * GAA(h) + ij];
            }
        A[r++] = tab; A[r++] = taa; A[r++] = tbb;
    } else {
        double t = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int h = SYM(i) ^ SYM(j);
                int ij = IAB(h, i, j);
                t += 2.0 * u[p->d2aboff[h] + (long)ij * GAB(h) + ij];
                if (i == j) continue;
                ij = IAA(h, i, j);
                t += u[p->d2aaoff[h] + (long)ij * GAA(h) + ij];
                t += u[p->d2bboff[h] + (long)ij * GAA(h) + ij];
            }
        A[r++] = t;
    }

    /* d2.cc:382-414 hermiticity: aa, bb, ab */
    for (int blk = 0; blk < 3; blk++)
        for (int h = 0; h < H; h++) {
            const long *off = blk == 0 ? p->d2aaoff : blk == 1 ? p->d2bboff : p->d2aboff;
            int g = blk == 2 ? GAB(h) : GAA(h);
            for (int ij = 0; ij < g; ij++)
                for (int kl = 0; kl < g; kl++)
                    A[r++] = u[off[h] + (long)ij * g + kl] - u[off[h] + (long)kl * g + ij];
        }

    /* d2.cc:416-434 D1(j,i) + Q1(i,j) */
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < H; h++) {
            int m = p->amopi[h];
            const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
            const long q1 = s == 0 ? p->q1aoff[h] : p->q1boff[h];
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++)
                    A[r + (long)i * m + j] = u[d1 + (long)j * m + i] + u[q1 + (long)i * m + j];
            r += (long)m * m;
        }

    /* d2.cc:442-592 contractions */
    if (!p->constrain_sz) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int m = p->amopi[h], o = p->pitzer_offset[h];
                const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
                const long *d2s = s == 0 ? p->d2aaoff : p->d2bboff;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        int ii = i + o, jj = j + o;
                        double acc = (na + nb - 1.0) * u[d1 + (long)i * m + j];
                        for (int k = 0; k < n; k++) {
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = s == 0 ? IAB(h2, ii, k) : IAB(h2, k, ii);
                            int jk = s == 0 ? IAB(h2, jj, k) : IAB(h2, k, jj);
                            acc -= u[p->d2aboff[h2] + (long)ik * GAB(h2) + jk];
                        }
                        for (int k = 0; k < n; k++) {
                            if (k == ii || k == jj) continue;
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = IAA(h2, ii, k), jk = IAA(h2, jj, k);
                            acc -= SGN(ii, k) * SGN(jj, k) * u[d2s[h2] + (long)ik * GAA(h2) + jk];
                        }
                        A[r + (long)i * m + j] = acc;
                    }
                r += (long)m * m;
            }
    } else {
        /* D2ab -> D1a, D2ab -> D1b */
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int m = p->amopi[h], o = p->pitzer_offset[h];
                const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
                double pre = s == 0 ? nb : na;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        int ii = i + o, jj = j + o;
                        double acc = pre * u[d1 + (long)i * m + j];
                        for (int k = 0; k < n; k++) {
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = s == 0 ? IAB(h2, ii, k) : IAB(h2, k, ii);
                            int jk = s == 0 ? IAB(h2, jj, k) : IAB(h2, k, jj);
                            acc -= u[p->d2aboff[h2] + (long)ik * GAB(h2) + jk];
                        }
                        A[r + (long)i * m + j] = acc;
                    }
                r += (long)m * m;
            }
        /* D2aa -> D1a, D2bb -> D1b */
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int m = p->amopi[h], o = p->pitzer_offset[h];
                const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
                const long *d2s = s == 0 ? p->d2aaoff : p->d2bboff;
                double pre = s == 0 ? na - 1.0 : nb - 1.0;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        int ii = i + o, jj = j + o;
                        double acc = pre * u[d1 + (long)i * m + j];
                        for (int k = 0; k < n; k++) {
                            if (k == ii || k == jj) continue;
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = IAA(h2, ii, k), jk = IAA(h2, jj, k);
                            acc -= SGN(ii, k) * SGN(jj, k) * u[d2s[h2] + (long)ik * GAA(h2) + jk];
                        }
                        A[r + (long)i * m + j] = acc;
                    }
                r += (long)m * m;
            }
    }

    /* spin.cc:276-508 */
    if (p->constrain_spin) {
        double s2 = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int h = SYM(i) ^ SYM(j);
                s2 += u[p->d2aboff[h] + (long)IAB(h, i, j) * GAB(h) + IAB(h, j, i)];
            }
        A[r++] = s2;
        if (p->nalpha == p->nbeta) {
            for (int h = 0; h < H; h++) {
                long mm = (long)p->amopi[h] * p->amopi[h];
                for (long t = 0; t < mm; t++) A[r + t] = u[p->d1aoff[h] + t] - u[p->d1boff[h] + t];
                r += mm;
            }
            for (int h = 0; h < H; h++) {
                long aa = (long)GAA(h) * GAA(h);
                for (long t = 0; t < aa; t++) A[r + t] = u[p->d2aaoff[h] + t] - u[p->d2bboff[h] + t];
                r += aa;
            }
            for (int s = 0; s < 2; s++) /* D2aa / D2bb from D2ab */
                for (int h = 0; h < H; h++) {
                    const long d2s = s == 0 ? p->d2aaoff[h] : p->d2bboff[h];
                    const long ab = p->d2aboff[h];
                    int ga = GAA(h), gb = GAB(h);
                    for (int ij = 0; ij < ga; ij++) {
                        int i = BAA0(h, ij), j = BAA1(h, ij);
                        long ijb = IAB(h, i, j), jib = IAB(h, j, i);
                        for (int kl = 0; kl < ga; kl++) {
                            int k = BAA0(h, kl), l = BAA1(h, kl);
                            long klb = IAB(h, k, l), lkb = IAB(h, l, k);
                            double v = u[d2s + (long)ij * ga + kl];
                            v -= 0.5 * u[ab + ijb * gb + klb];
                            v += 0.5 * u[ab + jib * gb + klb];
                            v += 0.5 * u[ab + ijb * gb + lkb];
                            v -= 0.5 * u[ab + jib * gb + lkb];
                            A[r + (long)ij * ga + kl] = v;
                        }
                    }
                    r += (long)ga * ga;
                }
            for (int h = 0; h < H; h++) { /* D200 */
                int g = GAB(h);
                const long ab = p->d2aboff[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    long ji = IAB(h, j, i);
                    double dij = i == j ? sqrt(2.0) : 1.0;
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long lk = IAB(h, l, k);
                        double dkl = k == l ? sqrt(2.0) : 1.0;
                        double f = 0.5 / (dij * dkl);
                        double v = u[p->d200off[h] + (long)ij * g + kl];
                        v -= f * u[ab + (long)ij * g + kl];
                        v -= f * u[ab + ji * g + kl];
                        v -= f * u[ab + (long)ij * g + lk];
                        v -= f * u[ab + ji * g + lk];
                        A[r + (long)ij * g + kl] = v;
                    }
                }
                r += (long)g * g;
            }
            for (int h = 0; h < H; h++) { /* D2ab(ij,kl) = D2ab(ji,lk) */
                int g = GAB(h);
                const long ab = p->d2aboff[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    long ji = IAB(h, j, i);
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long lk = IAB(h, l, k);
                        A[r + (long)ij * g + kl] = u[ab + (long)ij * g + kl] - u[ab + ji * g + lk];
                    }
                }
                r += (long)g * g;
            }
        } else {
            for (int h = 0; h < H; h++) {
                int g = GAB(h), g2 = 2 * GAB(h);
                const long ab = p->d2aboff[h], d0 = p->d200off[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    long ji = IAB(h, j, i);
                    double dij = i == j ? sqrt(2.0) : 1.0;
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long lk = IAB(h, l, k);
                        double dkl = k == l ? sqrt(2.0) : 1.0;
                        double xijkl = u[ab + (long)ij * g + kl], xjikl = u[ab + ji * g + kl];
                        double xijlk = u[ab + (long)ij * g + lk], xjilk = u[ab + ji * g + lk];
                        long q00 = (long)ij * g2 + kl, q01 = (long)ij * g2 + kl + g;
                        long q10 = (long)(ij + g) * g2 + kl, q11 = (long)(ij + g) * g2 + kl + g;
                        A[r + q00] = u[d0 + q00] - 0.5 / (dij * dkl) * (xijkl + xjikl + xijlk + xjilk);
                        A[r + q01] = u[d0 + q01] - 0.5 / dij * (xijkl - xijlk + xjikl - xjilk);
                        A[r + q10] = u[d0 + q10] - 0.5 / dkl * (xijkl + xijlk - xjikl - xjilk);
                        A[r + q11] = u[d0 + q11] - 0.5 * (xijkl - xjikl - xijlk + xjilk);
                    }
                }
                r += (long)g2 * g2;
            }
        }
        if (p->constrain_g2) {
            int g = GAB(0);
            for (int kl = 0; kl < g; kl++) {
                double acc = 0.0;
                for (int i = 0; i < n; i++) acc += u[p->g2baoff[0] + (long)kl * g + IAB(0, i, i)];
                A[r + kl] = acc;
            }
            r += g;
            for (int kl = 0; kl < g; kl++) {
                double acc = 0.0;
                for (int i = 0; i < n; i++) acc += u[p->g2baoff[0] + (long)IAB(0, i, i) * g + kl];
                A[r + kl] = acc;
            }
            r += g;
        }
    }

    /* q2.cc:170-317 */
    if (p->constrain_q2) {
        for (int h = 0; h < H; h++) {
            int g = GAB(h);
            for (int ij = 0; ij < g; ij++) {
                int i = BAB0(h, ij), j = BAB1(h, ij);
                for (int kl = 0; kl < g; kl++) {
                    int k = BAB0(h, kl), l = BAB1(h, kl);
                    long t = (long)ij * g + kl;
                    double v = u[p->d2aboff[h] + t] - u[p->q2aboff[h] + t];
                    if (j == l) v -= u[OB(p->d1aoff, k, i)];
                    if (i == k) v -= u[OB(p->d1boff, l, j)];
                    A[r + t] = v;
                }
            }
            r += (long)g * g;
        }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int g = GAA(h);
                const long d2 = s == 0 ? p->d2aaoff[h] : p->d2bboff[h];
                const long q2 = s == 0 ? p->q2aaoff[h] : p->q2bboff[h];
                const long *d1 = s == 0 ? p->d1aoff : p->d1boff;
                for (int ij = 0; ij < g; ij++) {
                    int i = BAA0(h, ij), j = BAA1(h, ij);
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAA0(h, kl), l = BAA1(h, kl);
                        long t = (long)ij * g + kl;
                        double v = u[d2 + t] - u[q2 + t];
                        if (j == l) v -= u[OB(d1, k, i)];
                        if (j == k) v += u[OB(d1, l, i)];
                        if (i == l) v += u[OB(d1, k, j)];
                        if (i == k) v -= u[OB(d1, l, j)];
                        A[r + t] = v;
                    }
                }
                r += (long)g * g;
            }
    }

    /* g2.cc:245-441 */
    if (p->constrain_g2) {
        for (int s = 0; s < 2; s++) /* G2ab, G2ba */
            for (int h = 0; h < H; h++) {
                int g = GAB(h);
                const long gof = s == 0 ? p->g2aboff[h] : p->g2baoff[h];
                const long *d1 = s == 0 ? p->d1aoff : p->d1boff;
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long t = (long)ij * g + kl;
                        double v = -u[gof + t];
                        if (j == l) v += u[OB(d1, i, k)];
                        int h2 = SYM(i) ^ SYM(l);
                        long a, bcol;
                        if (s == 0) { a = IAB(h2, i, l); bcol = IAB(h2, k, j); }
                        else        { a = IAB(h2, l, i); bcol = IAB(h2, j, k); }
                        v -= u[p->d2aboff[h2] + a * GAB(h2) + bcol];
                        A[r + t] = v;
                    }
                }
                r += (long)g * g;
            }
        for (int h = 0; h < H; h++) {
            int g = GAB(h), G = 2 * GAB(h);
            const long gof = p->g2aaoff[h];
            for (int ij = 0; ij < g; ij++) {
                int i = BAB0(h, ij), j = BAB1(h, ij);
                for (int kl = 0; kl < g; kl++) {
                    int k = BAB0(h, kl), l = BAB1(h, kl);
                    int h2 = SYM(i) ^ SYM(l);
                    for (int s = 0; s < 2; s++) { /* aaaa, bbbb */
                        long t = (long)(ij + s * g) * G + (kl + s * g);
                        double v = -u[gof + t];
                        if (j == l) v += u[OB(s == 0 ? p->d1aoff : p->d1boff, i, k)];
                        if (i != l && k != j) {
                            const long d2 = s == 0 ? p->d2aaoff[h2] : p->d2bboff[h2];
                            v -= SGN(i, l) * SGN(k, j) * u[d2 + (long)IAA(h2, i, l) * GAA(h2) + IAA(h2, k, j)];
                        }
                        A[r + t] = v;
                    }
                    long tab = (long)ij * G + (kl + g);      /* aabb */
                    A[r + tab] = -u[gof + tab] + u[p->d2aboff[h2] + (long)IAB(h2, i, l) * GAB(h2) + IAB(h2, j, k)];
                    long tba = (long)(ij + g) * G + kl;      /* bbaa */
                    A[r + tba] = -u[gof + tba] + u[p->d2aboff[h2] + (long)IAB(h2, l, i) * GAB(h2) + IAB(h2, k, j)];
                }
            }
            r += (long)G * G;
        }
    }
    if (p->constrain_t1) {
        struct t1_ctx c = {u, A};
        r = t1_rows(p, r, t1_gather, &c);
    }
    if (p->constrain_t2) {
        struct t1_ctx c = {u, A};
        r = t2_rows(p, r, t1_gather, &c);
    }
    if (p->constrain_d3) {
        struct t1_ctx c = {u, A};
        r = d3_rows(p, r, t1_gather, &c);
    }
}

/* ---- A^T u -------------------------------------------------------------- */
/* v2rdm_solver.cc:3052-3108; d2.cc:46-311, spin.cc:46-273, q2.cc:320-438, g2.cc:444-634 */
void orc_ATu(const orc_problem *p, const double *u, double *A) {
    const int n = NMO, H = p->nirrep;
    const double na = p->nalpha, nb = p->nbeta;
    memset(A, 0, sizeof(double) * p->n_primal);
    long r = 0;

    if (p->constrain_sz) {
        double yab = u[r++], yaa = u[r++], ybb = u[r++];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int h = SYM(i) ^ SYM(j);
                int ij = IAB(h, i, j);
                A[p->d2aboff[h] + (long)ij * GAB(h) + ij] += yab;
            }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                if (i == j) continue;
                int h = SYM(i) ^ SYM(j);
                int ij = IAA(h, i, j);
                A[p->d2aaoff[h] + (long)ij * GAA(h) + ij] += yaa;
                A[p->d2bboff[h] + (long)ij * GAA(h) + ij] += ybb;
            }
    } else {
        double y = u[r++];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int h = SYM(i) ^ SYM(j);
                int ij = IAB(h, i, j);
                A[p->d2aboff[h] + (long)ij * GAB(h) + ij] += 2.0 * y;
                if (i == j) continue;
                ij = IAA(h, i, j);
                A[p->d2aaoff[h] + (long)ij * GAA(h) + ij] += y;
                A[p->d2bboff[h] + (long)ij * GAA(h) + ij] += y;
            }
    }

    for (int blk = 0; blk < 3; blk++)
        for (int h = 0; h < H; h++) {
            const long *off = blk == 0 ? p->d2aaoff : blk == 1 ? p->d2bboff : p->d2aboff;
            int g = blk == 2 ? GAB(h) : GAA(h);
            for (int ij = 0; ij < g; ij++)
                for (int kl = 0; kl < g; kl++) {
                    double y = u[r++];
                    A[off[h] + (long)ij * g + kl] += y;
                    A[off[h] + (long)kl * g + ij] -= y;
                }
        }

    for (int s = 0; s < 2; s++)
        for (int h = 0; h < H; h++) {
            int m = p->amopi[h];
            const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
            const long q1 = s == 0 ? p->q1aoff[h] : p->q1boff[h];
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) {
                    double y = u[r + (long)i * m + j];
                    A[d1 + (long)j * m + i] += y;
                    A[q1 + (long)i * m + j] += y;
                }
            r += (long)m * m;
        }

    if (!p->constrain_sz) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int m = p->amopi[h], o = p->pitzer_offset[h];
                const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
                const long *d2s = s == 0 ? p->d2aaoff : p->d2bboff;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        int ii = i + o, jj = j + o;
                        double y = u[r + (long)i * m + j];
                        A[d1 + (long)i * m + j] += (na + nb - 1.0) * y;
                        for (int k = 0; k < n; k++) {
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = s == 0 ? IAB(h2, ii, k) : IAB(h2, k, ii);
                            int jk = s == 0 ? IAB(h2, jj, k) : IAB(h2, k, jj);
                            A[p->d2aboff[h2] + (long)ik * GAB(h2) + jk] -= y;
                        }
                        for (int k = 0; k < n; k++) {
                            if (k == ii || k == jj) continue;
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = IAA(h2, ii, k), jk = IAA(h2, jj, k);
                            A[d2s[h2] + (long)ik * GAA(h2) + jk] -= SGN(ii, k) * SGN(jj, k) * y;
                        }
                    }
                r += (long)m * m;
            }
    } else {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int m = p->amopi[h], o = p->pitzer_offset[h];
                const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
                double pre = s == 0 ? nb : na;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        int ii = i + o, jj = j + o;
                        double y = u[r + (long)i * m + j];
                        A[d1 + (long)i * m + j] += pre * y;
                        for (int k = 0; k < n; k++) {
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = s == 0 ? IAB(h2, ii, k) : IAB(h2, k, ii);
                            int jk = s == 0 ? IAB(h2, jj, k) : IAB(h2, k, jj);
                            A[p->d2aboff[h2] + (long)ik * GAB(h2) + jk] -= y;
                        }
                    }
                r += (long)m * m;
            }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int m = p->amopi[h], o = p->pitzer_offset[h];
                const long d1 = s == 0 ? p->d1aoff[h] : p->d1boff[h];
                const long *d2s = s == 0 ? p->d2aaoff : p->d2bboff;
                double pre = s == 0 ? na - 1.0 : nb - 1.0;
                for (int i = 0; i < m; i++)
                    for (int j = 0; j < m; j++) {
                        int ii = i + o, jj = j + o;
                        double y = u[r + (long)i * m + j];
                        A[d1 + (long)i * m + j] += pre * y;
                        for (int k = 0; k < n; k++) {
                            if (k == ii || k == jj) continue;
                            int h2 = SYM(ii) ^ SYM(k);
                            int ik = IAA(h2, ii, k), jk = IAA(h2, jj, k);
                            A[d2s[h2] + (long)ik * GAA(h2) + jk] -= SGN(ii, k) * SGN(jj, k) * y;
                        }
                    }
                r += (long)m * m;
            }
    }

    if (p->constrain_spin) {
        double y = u[r++];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                int h = SYM(i) ^ SYM(j);
                A[p->d2aboff[h] + (long)IAB(h, i, j) * GAB(h) + IAB(h, j, i)] += y;
            }
        if (p->nalpha == p->nbeta) {
            for (int h = 0; h < H; h++) {
                long mm = (long)p->amopi[h] * p->amopi[h];
                for (long t = 0; t < mm; t++) {
                    A[p->d1aoff[h] + t] += u[r + t];
                    A[p->d1boff[h] + t] -= u[r + t];
                }
                r += mm;
            }
            for (int h = 0; h < H; h++) {
                long aa = (long)GAA(h) * GAA(h);
                for (long t = 0; t < aa; t++) {
                    A[p->d2aaoff[h] + t] += u[r + t];
                    A[p->d2bboff[h] + t] -= u[r + t];
                }
                r += aa;
            }
            for (int s = 0; s < 2; s++)
                for (int h = 0; h < H; h++) {
                    const long d2s = s == 0 ? p->d2aaoff[h] : p->d2bboff[h];
                    const long ab = p->d2aboff[h];
                    int ga = GAA(h), gb = GAB(h);
                    for (int ij = 0; ij < ga; ij++) {
                        int i = BAA0(h, ij), j = BAA1(h, ij);
                        long ijb = IAB(h, i, j), jib = IAB(h, j, i);
                        for (int kl = 0; kl < ga; kl++) {
                            int k = BAA0(h, kl), l = BAA1(h, kl);
                            long klb = IAB(h, k, l), lkb = IAB(h, l, k);
                            double y = u[r + (long)ij * ga + kl];
                            A[d2s + (long)ij * ga + kl] += y;
                            A[ab + ijb * gb + klb] -= 0.5 * y;
                            A[ab + jib * gb + klb] += 0.5 * y;
                            A[ab + ijb * gb + lkb] += 0.5 * y;
                            A[ab + jib * gb + lkb] -= 0.5 * y;
                        }
                    }
                    r += (long)ga * ga;
                }
            for (int h = 0; h < H; h++) {
                int g = GAB(h);
                const long ab = p->d2aboff[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    long ji = IAB(h, j, i);
                    double dij = i == j ? sqrt(2.0) : 1.0;
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long lk = IAB(h, l, k);
                        double dkl = k == l ? sqrt(2.0) : 1.0;
                        double y = u[r + (long)ij * g + kl];
                        double f = 0.5 / (dij * dkl) * y;
                        A[p->d200off[h] + (long)ij * g + kl] += y;
                        A[ab + (long)ij * g + kl] -= f;
                        A[ab + ji * g + kl] -= f;
                        A[ab + (long)ij * g + lk] -= f;
                        A[ab + ji * g + lk] -= f;
                    }
                }
                r += (long)g * g;
            }
            for (int h = 0; h < H; h++) {
                int g = GAB(h);
                const long ab = p->d2aboff[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    long ji = IAB(h, j, i);
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long lk = IAB(h, l, k);
                        double y = u[r + (long)ij * g + kl];
                        A[ab + (long)ij * g + kl] += y;
                        A[ab + ji * g + lk] -= y;
                    }
                }
                r += (long)g * g;
            }
        } else {
            for (int h = 0; h < H; h++) {
                int g = GAB(h), g2 = 2 * GAB(h);
                const long ab = p->d2aboff[h], d0 = p->d200off[h];
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    long ji = IAB(h, j, i);
                    double dij = i == j ? sqrt(2.0) : 1.0;
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long lk = IAB(h, l, k);
                        double dkl = k == l ? sqrt(2.0) : 1.0;
                        long q00 = (long)ij * g2 + kl, q01 = (long)ij * g2 + kl + g;
                        long q10 = (long)(ij + g) * g2 + kl, q11 = (long)(ij + g) * g2 + kl + g;
                        double y00 = u[r + q00], y01 = u[r + q01], y10 = u[r + q10], y11 = u[r + q11];
                        A[d0 + q00] += y00; A[d0 + q01] += y01; A[d0 + q10] += y10; A[d0 + q11] += y11;
                        double f00 = 0.5 / (dij * dkl) * y00, f01 = 0.5 / dij * y01;
                        double f10 = 0.5 / dkl * y10, f11 = 0.5 * y11;
                        A[ab + (long)ij * g + kl] += -f00 - f01 - f10 - f11;
                        A[ab + ji * g + kl]       += -f00 - f01 + f10 + f11;
                        A[ab + (long)ij * g + lk] += -f00 + f01 - f10 + f11;
                        A[ab + ji * g + lk]       += -f00 + f01 + f10 - f11;
                    }
                }
                r += (long)g2 * g2;
            }
        }
        if (p->constrain_g2) {
            int g = GAB(0);
            for (int kl = 0; kl < g; kl++) {
                double y = u[r + kl];
                for (int i = 0; i < n; i++) A[p->g2baoff[0] + (long)kl * g + IAB(0, i, i)] += y;
            }
            r += g;
            for (int kl = 0; kl < g; kl++) {
                double y = u[r + kl];
                for (int i = 0; i < n; i++) A[p->g2baoff[0] + (long)IAB(0, i, i) * g + kl] += y;
            }
            r += g;
        }
    }

    if (p->constrain_q2) {
        for (int h = 0; h < H; h++) {
            int g = GAB(h);
            for (int ij = 0; ij < g; ij++) {
                int i = BAB0(h, ij), j = BAB1(h, ij);
                for (int kl = 0; kl < g; kl++) {
                    int k = BAB0(h, kl), l = BAB1(h, kl);
                    long t = (long)ij * g + kl;
                    double y = u[r + t];
                    A[p->d2aboff[h] + t] += y;
                    A[p->q2aboff[h] + t] -= y;
                    if (j == l) A[OB(p->d1aoff, k, i)] -= y;
                    if (i == k) A[OB(p->d1boff, l, j)] -= y;
                }
            }
            r += (long)g * g;
        }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int g = GAA(h);
                const long d2 = s == 0 ? p->d2aaoff[h] : p->d2bboff[h];
                const long q2 = s == 0 ? p->q2aaoff[h] : p->q2bboff[h];
                const long *d1 = s == 0 ? p->d1aoff : p->d1boff;
                for (int ij = 0; ij < g; ij++) {
                    int i = BAA0(h, ij), j = BAA1(h, ij);
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAA0(h, kl), l = BAA1(h, kl);
                        long t = (long)ij * g + kl;
                        double y = u[r + t];
                        A[d2 + t] += y;
                        A[q2 + t] -= y;
                        if (j == l) A[OB(d1, k, i)] -= y;
                        if (j == k) A[OB(d1, l, i)] += y;
                        if (i == l) A[OB(d1, k, j)] += y;
                        if (i == k) A[OB(d1, l, j)] -= y;
                    }
                }
                r += (long)g * g;
            }
    }

    if (p->constrain_g2) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int g = GAB(h);
                const long gof = s == 0 ? p->g2aboff[h] : p->g2baoff[h];
                const long *d1 = s == 0 ? p->d1aoff : p->d1boff;
                for (int ij = 0; ij < g; ij++) {
                    int i = BAB0(h, ij), j = BAB1(h, ij);
                    for (int kl = 0; kl < g; kl++) {
                        int k = BAB0(h, kl), l = BAB1(h, kl);
                        long t = (long)ij * g + kl;
                        double y = u[r + t];
                        A[gof + t] -= y;
                        if (j == l) A[OB(d1, i, k)] += y;
                        int h2 = SYM(i) ^ SYM(l);
                        long a, bcol;
                        if (s == 0) { a = IAB(h2, i, l); bcol = IAB(h2, k, j); }
                        else        { a = IAB(h2, l, i); bcol = IAB(h2, j, k); }
                        A[p->d2aboff[h2] + a * GAB(h2) + bcol] -= y;
                    }
                }
                r += (long)g * g;
            }
        for (int h = 0; h < H; h++) {
            int g = GAB(h), G = 2 * GAB(h);
            const long gof = p->g2aaoff[h];
            for (int ij = 0; ij < g; ij++) {
                int i = BAB0(h, ij), j = BAB1(h, ij);
                for (int kl = 0; kl < g; kl++) {
                    int k = BAB0(h, kl), l = BAB1(h, kl);
                    int h2 = SYM(i) ^ SYM(l);
                    for (int s = 0; s < 2; s++) {
                        long t = (long)(ij + s * g) * G + (kl + s * g);
                        double y = u[r + t];
                        A[gof + t] -= y;
                        if (j == l) A[OB(s == 0 ? p->d1aoff : p->d1boff, i, k)] += y;
                        if (i != l && k != j) {
                            const long d2 = s == 0 ? p->d2aaoff[h2] : p->d2bboff[h2];
                            A[d2 + (long)IAA(h2, i, l) * GAA(h2) + IAA(h2, k, j)] -= SGN(i, l) * SGN(k, j) * y;
                        }
                    }
                    long tab = (long)ij * G + (kl + g);
                    A[gof + tab] -= u[r + tab];
                    A[p->d2aboff[h2] + (long)IAB(h2, i, l) * GAB(h2) + IAB(h2, j, k)] += u[r + tab];
                    long tba = (long)(ij + g) * G + kl;
                    A[gof + tba] -= u[r + tba];
                    A[p->d2aboff[h2] + (long)IAB(h2, l, i) * GAB(h2) + IAB(h2, k, j)] += u[r + tba];
                }
            }
            r += (long)G * G;
        }
    }
    if (p->constrain_t1) {
        struct t1_ctx c = {u, A};
        r = t1_rows(p, r, t1_scatter, &c);
    }
    if (p->constrain_t2) {
        struct t1_ctx c = {u, A};
        r = t2_rows(p, r, t1_scatter, &c);
    }
    if (p->constrain_d3) {
        struct t1_ctx c = {u, A};
        r = d3_rows(p, r, t1_scatter, &c);
    }
}

void orc_cb_Au(double *out, const double *in, void *data) { orc_Au((const orc_problem *)data, in, out); }
void orc_cb_ATu(double *out, const double *in, void *data) { orc_ATu((const orc_problem *)data, in, out); }

/* ---- guess / post-processing --------------------------------------------- */

/* v2rdm_solver.cc:1799-1806 */
void orc_guess_random(long n, double *x) {
    srand(0);
    for (long i = 0; i < n; i++) x[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
}

/* v2rdm_solver.cc:1370-1381: returns sum_ij D2ab(ij,ji); S^2 = N/2 + ms^2 - that */
double orc_spin_squared(const orc_problem *p, const double *x) {
    int n = NMO;
    double s = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            int h = SYM(i) ^ SYM(j);
            s += x[p->d2aboff[h] + (long)IAB(h, i, j) * GAB(h) + IAB(h, j, i)];
        }
    double ms = (p->multiplicity - 1.0) / 2.0;
    return 0.5 * (p->nalpha + p->nbeta) + ms * ms - s;
}

/* python_api.cc:374-394, as a dense n x n matrix in Pitzer order */
void orc_opdm(const orc_problem *p, const double *x, double *d) {
    int n = NMO;
    memset(d, 0, sizeof(double) * n * n);
    for (int h = 0; h < p->nirrep; h++) {
        int m = p->amopi[h], o = p->pitzer_offset[h];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++)
                d[(long)(i + o) * n + (j + o)] = x[p->d1aoff[h] + (long)i * m + j] + x[p->d1boff[h] + (long)i * m + j];
    }
}

/* python_api.cc:396-456: tpdm[(i*n+k)*n*n + j*n+l] = D2ab(ij,kl)+D2ab(ji,lk)+s(D2aa+D2bb)(ij,kl) */
int orc_tpdm(const orc_problem *p, const double *x, double *t) {
    if (p->nirrep > 1) return -1;
    int n = NMO;
    long g = GAB(0), ga = GAA(0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++)
                for (int l = 0; l < n; l++) {
                    double v = x[p->d2aboff[0] + IAB(0, i, j) * g + IAB(0, k, l)]
                             + x[p->d2aboff[0] + IAB(0, j, i) * g + IAB(0, l, k)];
                    if (i != j && k != l) {
                        double sg = ((i > j) ? -1.0 : 1.0) * ((k > l) ? -1.0 : 1.0);
                        long e = IAA(0, i, j) * ga + IAA(0, k, l);
                        v += sg * (x[p->d2aaoff[0] + e] + x[p->d2bboff[0] + e]);
                    }
                    t[((long)i * n + k) * n * n + (long)j * n + l] = v;
                }
    return 0;
}

/* ---- projection ---------------------------------------------------------- */

/* Cyclic Jacobi on a symmetric d x d matrix (row-major, overwritten).
 * Eigenvalues in w, eigenvectors in the columns of v. */
void orc_jacobi_eig(int d, double *a, double *w, double *v) {
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) v[(long)i * d + j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < d; i++)
            for (int j = 0; j < d; j++) {
                double q = a[(long)i * d + j] * a[(long)i * d + j];
                tot += q;
                if (i != j) off += q;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int pp = 0; pp < d - 1; pp++)
            for (int q = pp + 1; q < d; q++) {
                double apq = a[(long)pp * d + q];
                if (apq == 0.0) continue;
                double app = a[(long)pp * d + pp], aqq = a[(long)q * d + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < d; k++) {
                    double akp = a[(long)k * d + pp], akq = a[(long)k * d + q];
                    a[(long)k * d + pp] = c * akp - s * akq;
                    a[(long)k * d + q] = s * akp + c * akq;
                }
                for (int k = 0; k < d; k++) {
                    double apk = a[(long)pp * d + k], aqk = a[(long)q * d + k];
                    a[(long)pp * d + k] = c * apk - s * aqk;
                    a[(long)q * d + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < d; k++) {
                    double vkp = v[(long)k * d + pp], vkq = v[(long)k * d + q];
                    v[(long)k * d + pp] = c * vkp - s * vkq;
                    v[(long)k * d + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < d; i++) w[i] = a[(long)i * d + i];
}

/* U (d x d) -> x = V max(L,0) V^T / mu, z = -V min(L,0) V^T, after
 * symmetrising U.  work holds at least 3 d*d doubles (+ LAPACK workspace). */
int orc_project_block(int d, const double *U, double mu, double *x, double *z,
                      orc_dsyevd_fn eig, double *work, int lwork, int *iwork, int liwork) {
    if (d == 0) return 0;
    long dd = (long)d * d;
    double *a = work, *w = work + dd, *v = work + dd + d, *lw = work + 2 * dd + d;
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) a[(long)i * d + j] = 0.5 * (U[(long)i * d + j] + U[(long)j * d + i]);
    if (eig) {
        int info = 0, n = d, lda = d;
        int lw_avail = lwork - (int)(2 * dd + d);
        eig("V", "U", &n, a, &lda, w, lw, &lw_avail, iwork, &liwork, &info, 1, 1);
        if (info != 0) return info;
        /* column-major result: eigenvector k is a[k*d + i] */
        for (int i = 0; i < d; i++)
            for (int k = 0; k < d; k++) v[(long)i * d + k] = a[(long)k * d + i];
    } else {
        orc_jacobi_eig(d, a, w, v);
    }
    for (int i = 0; i < d; i++)
        for (int j = i; j < d; j++) {
            double sp = 0.0, sn = 0.0;
            for (int k = 0; k < d; k++) {
                double t = v[(long)i * d + k] * v[(long)j * d + k] * w[k];
                if (w[k] > 0.0) sp += t; else sn += t;
            }
            x[(long)i * d + j] = x[(long)j * d + i] = sp / mu;
            z[(long)i * d + j] = z[(long)j * d + i] = -sn;
        }
    return 0;
}

/* ---- BPSDP ----------------------------------------------------------------- */

void orc_sdp_state_init(orc_sdp_state *s, long np, long nd) {
    memset(s, 0, sizeof(*s));
    s->n_primal = np;
    s->n_dual = nd;
    s->y = (double *)calloc(nd, sizeof(double));
    s->z = (double *)calloc(np, sizeof(double));
}

void orc_sdp_state_free(orc_sdp_state *s) {
    free(s->y);
    free(s->z);
    memset(s, 0, sizeof(*s));
}

static double dot(long n, const double *a, const double *b) {
    double s = 0.0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Relative CG residual tolerance for one outer iteration:
 *   ||r|| <= tau ||rhs||,  tau = clamp(0.003 max(eps_p, eps_d), CG_CONVERGENCE, 1e-2),
 * with eps = 1 before the first outer iteration.  (libsdp's schedule is not
 * available; this one was chosen on the 8-site Hubbard ring, where it gave
 * the fewest outer and CG iterations among the factors tried.) */
double orc_cg_tolerance(const orc_sdp_options *o, long oiter, double eps_p, double eps_d) {
    double e = eps_p > eps_d ? eps_p : eps_d;
    if (oiter == 0) e = 1.0;
    double tol = 0.003 * e;
    if (tol > 1e-2) tol = 1e-2;
    if (tol < o->cg_convergence) tol = o->cg_convergence;
    return tol;
}

/*
 * Boundary-point method (Povh, Rendl, Wiegele, Computing 78, 277 (2006);
 * Malick, Povh, Rendl, Wiegele, SIAM J. Optim. 20, 336 (2009); Mazziotti,
 * PRL 106, 083001 (2011)), the iteration libsdp's BPSDPSolver::solve runs for
 * v2rdm_solver.cc:1340.  Per outer iteration:
 *   1. solve (A A^T) y = A(c - z) + mu (b - A x) by CG, warm-started at y;
 *   2. U = mu x + A^T y - c; per block U = V L V^T (symmetrised);
 *      x <- V L+ V^T / mu,  z <- -V L- V^T;
 *   3. eps_p = ||A x - b|| / ||b||, eps_d = ||A^T y - c + z|| / ||c||,
 *      E_p = c.x, E_d = b.y;
 *   4. every MU_UPDATE_FREQUENCY outer iterations mu <- mu * eps_p / eps_d;
 *   5. converged when |E_p - E_d| < E_CONVERGENCE and eps_p, eps_d < R_CONVERGENCE.
 */
int orc_bpsdp_solve(orc_sdp_state *s, const orc_sdp_options *o,
                    double *x, const double *b, const double *c,
                    const int *dims, int nblocks, int maxiter,
                    orc_map_fn Au, orc_map_fn ATu, void *data,
                    orc_dsyevd_fn eig, orc_monitor_fn mon, int print_level, void *mon_data) {
    long np = s->n_primal, nd = s->n_dual;
    if (!s->initialized) {
        memset(s->y, 0, sizeof(double) * nd);
        memset(s->z, 0, sizeof(double) * np);
        s->mu = o->mu0 > 0 ? o->mu0 : 1.0;
        s->eps_p = s->eps_d = 1.0;
        s->initialized = 1;
    }
    double *Ax = (double *)malloc(sizeof(double) * nd);
    double *rhs = (double *)malloc(sizeof(double) * nd);
    double *r = (double *)malloc(sizeof(double) * nd);
    double *pv = (double *)malloc(sizeof(double) * nd);
    double *w = (double *)malloc(sizeof(double) * nd);
    double *tp = (double *)malloc(sizeof(double) * np);
    double *ATy = (double *)malloc(sizeof(double) * np);
    double *U = (double *)malloc(sizeof(double) * np);

    double bn = sqrt(dot(nd, b, b)), cn = sqrt(dot(np, c, c));
    if (bn == 0.0) bn = 1.0;
    if (cn == 0.0) cn = 1.0;

    long boff[14 * ORC_MAX_IRREP + 1];
    int maxd = 0;
    boff[0] = 0;
    for (int k = 0; k < nblocks; k++) {
        boff[k + 1] = boff[k] + (long)dims[k] * dims[k];
        if (dims[k] > maxd) maxd = dims[k];
    }
    int lwork = 2 * maxd * maxd + maxd + 1 + 6 * maxd + 2 * maxd * maxd + 64;
    int liwork = 3 + 5 * maxd + 16;

    s->converged = 0;
    Au(Ax, x, data);
    for (int it = 0; it < maxiter; it++) {
        if (o->maxiter > 0 && s->oiter_total >= o->maxiter) break;
        double t0 = wall();
        /* rhs = A(c - z) + mu (b - Ax) */
        for (long i = 0; i < np; i++) tp[i] = c[i] - s->z[i];
        Au(rhs, tp, data);
        for (long i = 0; i < nd; i++) rhs[i] += s->mu * (b[i] - Ax[i]);
        /* r = rhs - A A^T y */
        ATu(tp, s->y, data);
        Au(w, tp, data);
        for (long i = 0; i < nd; i++) { r[i] = rhs[i] - w[i]; pv[i] = r[i]; }
        double rr = dot(nd, r, r);
        double tol = orc_cg_tolerance(o, s->oiter_total, s->eps_p, s->eps_d);
        double thr = tol * tol * dot(nd, rhs, rhs);
        int k = 0;
        while (k < o->cg_maxiter && rr > thr) {
            ATu(tp, pv, data);
            Au(w, tp, data);
            double pw = dot(nd, pv, w);
            if (pw <= 0.0) break;
            double alpha = rr / pw;
            for (long i = 0; i < nd; i++) { s->y[i] += alpha * pv[i]; r[i] -= alpha * w[i]; }
            double rr1 = dot(nd, r, r);
            double beta = rr1 / rr;
            rr = rr1;
            for (long i = 0; i < nd; i++) pv[i] = r[i] + beta * pv[i];
            k++;
        }
        s->iiter_total += k;
        double t1 = wall();
        s->cg_time += t1 - t0;

        /* U = mu x + A^T y - c, project each block */
        ATu(ATy, s->y, data);
        for (long i = 0; i < np; i++) U[i] = s->mu * x[i] + ATy[i] - c[i];
        int fail = 0;
#pragma omp parallel num_threads(o->nthreads > 0 ? o->nthreads : 1) reduction(+ : fail)
        {
            double *work = (double *)malloc(sizeof(double) * (lwork + 3L * maxd * maxd));
            int *iwork = (int *)malloc(sizeof(int) * liwork);
#pragma omp for schedule(dynamic, 1)
            for (int kb = 0; kb < nblocks; kb++) {
                if (orc_project_block(dims[kb], U + boff[kb], s->mu, x + boff[kb], s->z + boff[kb],
                                      eig, work, lwork + 3 * maxd * maxd, iwork, liwork) != 0)
                    fail += 1;
            }
            free(work);
            free(iwork);
        }
        double t2 = wall();
        s->proj_time += t2 - t1;
        if (fail) break;

        /* errors and energies */
        double ed = 0.0;
        for (long i = 0; i < np; i++) { double t = ATy[i] - c[i] + s->z[i]; ed += t * t; }
        Au(Ax, x, data);
        double ep = 0.0;
        for (long i = 0; i < nd; i++) { double t = Ax[i] - b[i]; ep += t * t; }
        s->eps_p = sqrt(ep) / bn;
        s->eps_d = sqrt(ed) / cn;
        s->energy_primal = dot(np, c, x);
        s->energy_dual = dot(nd, b, s->y);
        s->oiter_total++;
        if (o->mu_update_frequency > 0 && s->oiter_total % o->mu_update_frequency == 0 &&
            s->eps_d > 0.0 && s->eps_p > 0.0)
            s->mu *= s->eps_p / s->eps_d;
        s->cg_time += wall() - t2;
        if (mon) mon(print_level, (int)s->oiter_total, k, s->energy_primal, s->energy_dual, s->mu,
                     s->eps_p, s->eps_d, mon_data);
        if (fabs(s->energy_primal - s->energy_dual) < o->e_convergence &&
            s->eps_p < o->r_convergence && s->eps_d < o->r_convergence) {
            s->converged = 1;
            break;
        }
    }
    free(Ax); free(rhs); free(r); free(pv); free(w); free(tp); free(ATy); free(U);
    return s->converged;
}

/* ---- heap constructors for ctypes callers -------------------------------- */
orc_problem *orc_problem_new(int nmo, int nalpha, int nbeta, int multiplicity,
                             const int *irreps, int nirrep, int q2, int g2, int spin, int sz) {
    orc_problem *p = (orc_problem *)calloc(1, sizeof(orc_problem));
    if (orc_problem_init(p, nmo, nalpha, nbeta, multiplicity, irreps, nirrep, q2, g2, spin, sz) != 0) {
        free(p);
        return NULL;
    }
    return p;
}
void orc_problem_delete(orc_problem *p) {
    if (!p) return;
    orc_problem_free(p);
    free(p);
}
orc_sdp_state *orc_sdp_state_new(long np, long nd) {
    orc_sdp_state *s = (orc_sdp_state *)calloc(1, sizeof(orc_sdp_state));
    orc_sdp_state_init(s, np, nd);
    return s;
}
void orc_sdp_state_delete(orc_sdp_state *s) {
    if (!s) return;
    orc_sdp_state_free(s);
    free(s);
}
/* scalars of a state: mu, E_p, E_d, eps_p, eps_d, cg_time, proj_time */
void orc_sdp_state_scalars(const orc_sdp_state *s, double *out, long *counts) {
    out[0] = s->mu; out[1] = s->energy_primal; out[2] = s->energy_dual;
    out[3] = s->eps_p; out[4] = s->eps_d; out[5] = s->cg_time; out[6] = s->proj_time;
    counts[0] = s->oiter_total; counts[1] = s->iiter_total; counts[2] = s->converged;
}
double *orc_sdp_state_y(orc_sdp_state *s) { return s->y; }
double *orc_sdp_state_z(orc_sdp_state *s) { return s->z; }
/* Solve with the oracle's own DQG maps. */
int orc_solve_dqg(orc_sdp_state *s, const orc_sdp_options *o, const orc_problem *p,
                  double *x, const double *b, const double *c, int maxiter,
                  orc_dsyevd_fn eig, orc_monitor_fn mon, int print_level) {
    return orc_bpsdp_solve(s, o, x, b, c, p->dims, p->nblocks, maxiter, orc_cb_Au, orc_cb_ATu,
                           (void *)p, eig, mon, print_level, NULL);
}
