/*
 * v2rdm_oracle.h -- FP64 CPU oracle for the fixed-orbital v2RDM DQG solve.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (hilbert_amd) never links,
 * loads or calls it.
 *
 * This is a plain-loop restatement of edeprince3/hilbert's v2RDM-CASSCF
 * index tables, b/c packing and DQG+spin constraint maps, plus the published
 * boundary-point SDP iteration that libsdp implements.  Each function cites
 * the reference file:line it follows.
 *
 * Pinning (see tests/test_oracle_*.py): adjointness <y,Au> = <A^T y,u>;
 * exact-RDM feasibility A x(Psi) = b for brute-force wavefunctions;
 * c.x(Psi) = <Psi|H|Psi>; closed forms (2-site Hubbard, U = 0, two
 * electrons); E_DQG <= E_FCI; symmetry-blocked == C1 solve; table invariants.
 */
#ifndef V2RDM_ORACLE_H
#define V2RDM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_IRREP 8

typedef struct orc_problem {
    /* inputs */
    int nmo;                 /* active orbitals (amo_) */
    int nirrep;
    int nalpha, nbeta;       /* active electrons */
    int multiplicity;
    int constrain_q2, constrain_g2, constrain_spin, constrain_sz;
    int *symmetry;           /* [nmo] irrep of each orbital, Pitzer order */

    /* derived (basis.cc:104-111, 413-419, 427-533) */
    int amopi[ORC_MAX_IRREP];
    int pitzer_offset[ORC_MAX_IRREP];
    int gems_ab[ORC_MAX_IRREP];
    int gems_aa[ORC_MAX_IRREP];
    int *bas_ab[ORC_MAX_IRREP];   /* [gems_ab*2]: (first, second) */
    int *ibas_ab[ORC_MAX_IRREP];  /* [nmo*nmo], -999 when the pair is not in irrep h */
    int *bas_aa[ORC_MAX_IRREP];   /* [gems_aa*2] */
    int *ibas_aa[ORC_MAX_IRREP];  /* [nmo*nmo] */

    /* primal offsets (v2rdm_solver.cc:3729-3846) */
    long d2aboff[ORC_MAX_IRREP], d2aaoff[ORC_MAX_IRREP], d2bboff[ORC_MAX_IRREP], d200off[ORC_MAX_IRREP];
    long d1aoff[ORC_MAX_IRREP], d1boff[ORC_MAX_IRREP], q1aoff[ORC_MAX_IRREP], q1boff[ORC_MAX_IRREP];
    long q2aboff[ORC_MAX_IRREP], q2aaoff[ORC_MAX_IRREP], q2bboff[ORC_MAX_IRREP];
    long g2aboff[ORC_MAX_IRREP], g2baoff[ORC_MAX_IRREP], g2aaoff[ORC_MAX_IRREP];
    long n_primal, n_dual;
    int nblocks;
    int dims[26 * ORC_MAX_IRREP];

    /* T1 / T2 (POSITIVITY DQGT1, DQGT2, DQGT1T2), set by orc_problem_enable_t:
     * triplet tables (basis.cc:642-744), the T1 blocks after G2
     * (v2rdm_solver.cc:3848-3873) and the T2 blocks after them (:3879-3908) */
    int constrain_t1, constrain_t2, constrain_d3;
    int trip_aab[ORC_MAX_IRREP], trip_aaa[ORC_MAX_IRREP], trip_aba[ORC_MAX_IRREP];
    int *bas_aab[ORC_MAX_IRREP], *bas_aaa[ORC_MAX_IRREP], *bas_aba[ORC_MAX_IRREP];  /* [trip*3] */
    int *ibas_aab[ORC_MAX_IRREP], *ibas_aaa[ORC_MAX_IRREP];  /* [nmo^3], -999 when absent */
    long t1aaaoff[ORC_MAX_IRREP], t1bbboff[ORC_MAX_IRREP], t1aaboff[ORC_MAX_IRREP], t1bbaoff[ORC_MAX_IRREP];
    long t2aaaoff[ORC_MAX_IRREP], t2bbboff[ORC_MAX_IRREP], t2aaboff[ORC_MAX_IRREP], t2bbaoff[ORC_MAX_IRREP];
    long d3aaaoff[ORC_MAX_IRREP], d3bbboff[ORC_MAX_IRREP], d3aaboff[ORC_MAX_IRREP], d3bbaoff[ORC_MAX_IRREP];
} orc_problem;

/* Build tables/offsets.  irreps: [nmo] orbital irreps in Pitzer (non-decreasing)
 * order; nirrep in {1,2,4,8}.  Returns 0 on success. */
int  orc_problem_init(orc_problem *p, int nmo, int nalpha, int nbeta, int multiplicity,
                      const int *irreps, int nirrep,
                      int constrain_q2, int constrain_g2, int constrain_spin, int constrain_sz);
void orc_problem_free(orc_problem *p);
/* Add the T1 (t1.cc), T2 (t2.cc, the slow path used at v2rdm_solver.cc:3021)
 * and D3 (d3.cc, CONSTRAIN_D3) conditions to an initialised problem:
 * triplet tables, the T1aaa, T1bbb, T1aab, T1bba blocks, then the T2aaa,
 * T2bbb (dimension trip_aab + trip_aba), T2aab, T2bba blocks, then the D3aaa,
 * D3bbb, D3aab, D3bba blocks, and their rows (b = 0, v2rdm_solver.cc:
 * 2076-2150, 2269-2380).  Call once.  Returns 0, or -1 when T1 or T2 is
 * asked for without Q2 and G2. */
int  orc_problem_enable_t(orc_problem *p, int t1, int t2, int d3);
/* [12][nirrep] offsets of the T1aaa, T1bbb, T1aab, T1bba, T2aaa, T2bbb,
 * T2aab, T2bba, D3aaa, D3bbb, D3aab, D3bba blocks */
void orc_get_t_offsets(const orc_problem *p, long *off);
void orc_get_trip_counts(const orc_problem *p, int *trip_aab, int *trip_aaa, int *trip_aba);

/* Accessors for ctypes. */
long orc_n_primal(const orc_problem *p);
long orc_n_dual(const orc_problem *p);
int  orc_nblocks(const orc_problem *p);
void orc_get_dims(const orc_problem *p, int *dims);
void orc_get_gems(const orc_problem *p, int *gems_ab, int *gems_aa);
/* Dense table export: bas[h] rows laid out contiguously, ibas as [nirrep][nmo][nmo]. */
void orc_get_tables(const orc_problem *p, int *bas_ab, int *ibas_ab, int *bas_aa, int *ibas_aa);
/* 14 offset arrays, [14][nirrep] in the primal block order. */
void orc_get_offsets(const orc_problem *p, long *off);

/* b vector (v2rdm_solver.cc:1826-2074). */
void orc_build_b(const orc_problem *p, double *b);

/* c vector from h[n*n], g[n^4] (chemist (pq|rs) at g[((p*n+q)*n+r)*n+s]), in
 * Pitzer order (v2rdm_solver.cc:164-193, tei.cc:139-199, 201-275). */
void orc_pack_c(const orc_problem *p, const double *h, const double *g, double *c);
/* Adjoint of orc_pack_c: hbar[n*n], gbar[n^4] (overwritten). */
void orc_pack_c_adjoint(const orc_problem *p, const double *xbar, double *hbar, double *gbar);
/* Hubbard ring integrals (oei.cc:76-95, tei.cc:285-291). */
void orc_hubbard_integrals(int n, double t, double U, double *h, double *g);

/* Density-fitted Hamiltonian of the active space with the frozen-core terms:
 * the DF branch of TEI (tei.cc:280-283, TEI(i,j,k,l) = sum_Q
 * qmo[Q*ntri + INDEX(i,j)] qmo[Q*ntri + INDEX(k,l)], ntri = nfull(nfull+1)/2,
 * INDEX the packed lower triangle), FrozenCoreEnergy (tei.cc:201-275) and the
 * integrals RepackIntegrals reads (tei.cc:139-199).  h_full[nfull*nfull] is
 * read on and above the diagonal, as oei_full_sym_ holds it (tei.cc:124-132).
 * core[ncore] and active[nact] are full-space orbital indices; the outputs
 * are in the order of active[]: h_act[nact*nact] (with the core Fock terms),
 * g_act[nact^4] chemists' (pq|rs), *efzc the frozen-core energy. */
void orc_df_integrals(int nfull, const double *h_full, int nq, const double *qmo, int ncore, const int *core,
                      int nact, const int *active, double *h_act, double *g_act, double *efzc);

/* FCIDUMP records (fcidump.cc:60-243), in memory: one record is a value and
 * four orbital labels, in the order the reference fwrite()s them.
 * orc_fcidump_map: the label of every orbital, 1-based, ordered by space
 * (space[p]: 0 core, 1 active, 2 virtual), then irrep, then index
 * (fcidump.cc:83-107); full_irreps may be NULL (C1).
 * orc_fcidump_integrals: the (pq|rs) with p <= q, r <= s, INDEX(p,q) <=
 * INDEX(r,s) of equal pair symmetry from the DF integrals, then h[p][q] over
 * each irrep block, then (enuc, 0, 0, 0, 0) (fcidump.cc:109-168).
 * orc_fcidump_rdms: the spin-free 2-RDM over the ab geminals, then the
 * spin-summed 1-RDM (fcidump.cc:170-243); act_full[i] is the full-space index
 * of Pitzer active orbital i, labels are map[] - ninact.
 * Both return the record count; with val == NULL they only count. */
void orc_fcidump_map(int nfull, int nirrep, const int *full_irreps, const int *space, int *map);
long orc_fcidump_integrals(int nfull, const int *full_irreps, const int *map, const double *h_full, int nq,
                           const double *qmo, double enuc, double *val, int *lab);
long orc_fcidump_rdms(const orc_problem *p, const double *x, const int *act_full, const int *map, int ninact,
                      double *val, int *lab);

/* Constraint maps (v2rdm_solver.cc:2995-3108; d2.cc, spin.cc, q2.cc, g2.cc). */
void orc_Au(const orc_problem *p, const double *u, double *out);
void orc_ATu(const orc_problem *p, const double *u, double *out);

/* Random guess (v2rdm_solver.cc:1799-1806): srand(0), glibc rand(). */
void orc_guess_random(long n, double *x);

/* Post-processing (v2rdm_solver.cc:1370-1397, python_api.cc:374-456). */
double orc_spin_squared(const orc_problem *p, const double *x);
void   orc_opdm(const orc_problem *p, const double *x, double *opdm);      /* [n*n] D1a+D1b, Pitzer */
int    orc_tpdm(const orc_problem *p, const double *x, double *tpdm);      /* [n^4], C1 only */

/* ---------------- boundary-point SDP (published method) ---------------- */
typedef void (*orc_map_fn)(double *out, const double *in, void *data);
typedef void (*orc_dsyevd_fn)(const char *jobz, const char *uplo, const int *n, double *a,
                              const int *lda, double *w, double *work, const int *lwork,
                              int *iwork, const int *liwork, int *info,
                              unsigned long jobz_len, unsigned long uplo_len);
typedef void (*orc_monitor_fn)(int print_level, int oiter, int iiter, double ep, double ed,
                               double mu, double eps_p, double eps_d, void *data);

typedef struct orc_sdp_options {
    double e_convergence;        /* E_CONVERGENCE */
    double r_convergence;        /* R_CONVERGENCE */
    double cg_convergence;       /* CG_CONVERGENCE: floor of the relative CG tolerance */
    int    cg_maxiter;           /* CG_MAXITER */
    int    maxiter;              /* MAXITER: cap on total outer iterations */
    int    mu_update_frequency;  /* MU_UPDATE_FREQUENCY */
    double mu0;                  /* initial penalty parameter */
    int    nthreads;             /* OpenMP threads for the projection */
} orc_sdp_options;

typedef struct orc_sdp_state {
    long n_primal, n_dual;
    double *y, *z;
    double mu;
    int initialized, converged;
    long oiter_total, iiter_total;
    double energy_primal, energy_dual, eps_p, eps_d;
    double cg_time, proj_time;
} orc_sdp_state;

void orc_sdp_state_init(orc_sdp_state *s, long n_primal, long n_dual);
void orc_sdp_state_free(orc_sdp_state *s);

/* One call = up to maxiter outer iterations (libsdp SDPSolver::solve contract,
 * v2rdm_solver.cc:1340).  y, z, mu persist in the state across calls.
 * eig may be NULL (then a cyclic Jacobi fallback is used).
 * Returns 1 if converged, 0 otherwise. */
int orc_bpsdp_solve(orc_sdp_state *s, const orc_sdp_options *o,
                    double *x, const double *b, const double *c,
                    const int *dims, int nblocks, int maxiter,
                    orc_map_fn Au, orc_map_fn ATu, void *data,
                    orc_dsyevd_fn eig, orc_monitor_fn mon, int print_level, void *mon_data);

/* Convenience: the DQG maps of an orc_problem as orc_map_fn callbacks. */
void orc_cb_Au(double *out, const double *in, void *data);
void orc_cb_ATu(double *out, const double *in, void *data);

/* CG tolerance policy shared with the device solver (see orc_bpsdp_solve). */
double orc_cg_tolerance(const orc_sdp_options *o, long oiter, double eps_p, double eps_d);

/* PSD projection of one symmetric block (symmetrises first): x = U+/mu, z = -U-.
 * Exposed for tests. */
int orc_project_block(int d, const double *U, double mu, double *x, double *z,
                      orc_dsyevd_fn eig, double *work, int lwork, int *iwork, int liwork);
/* Eigen-decomposition by cyclic Jacobi (fallback, and an independent check). */
void orc_jacobi_eig(int d, double *a, double *w, double *v);

/* Heap constructors and accessors for ctypes callers. */
orc_problem   *orc_problem_new(int nmo, int nalpha, int nbeta, int multiplicity,
                               const int *irreps, int nirrep, int q2, int g2, int spin, int sz);
void           orc_problem_delete(orc_problem *p);
orc_sdp_state *orc_sdp_state_new(long n_primal, long n_dual);
void           orc_sdp_state_delete(orc_sdp_state *s);
void           orc_sdp_state_scalars(const orc_sdp_state *s, double *out, long *counts);
double        *orc_sdp_state_y(orc_sdp_state *s);
double        *orc_sdp_state_z(orc_sdp_state *s);
int            orc_solve_dqg(orc_sdp_state *s, const orc_sdp_options *o, const orc_problem *p,
                             double *x, const double *b, const double *c, int maxiter,
                             orc_dsyevd_fn eig, orc_monitor_fn mon, int print_level);

#ifdef __cplusplus
}
#endif
#endif
