/*
 * v2rdm_hip.h -- C ABI of the MI355X-native fixed-orbital v2RDM (DQG) solver.
 *
 * The boundary replaces, in edeprince3/hilbert:
 *   - libsdp::BPSDPSolver (constructed at src/v2rdm_casscf/v2rdm_solver.cc:424-433,
 *     driven by sdp_->solve(...) at v2rdm_solver.cc:1340 and the
 *     do { solve } while (!is_converged()) loop at v2rdm_solver.cc:1330-1361),
 *   - the callbacks it calls back into, v2RDMSolver::bpsdp_Au / bpsdp_ATu
 *     (v2rdm_solver.cc:2995-3108, trampolines evaluate_Au / evaluate_ATu at
 *     v2rdm_solver.cc:97-112),
 *   - the setup they rely on: build_mapping_arrays (basis.cc:44-535),
 *     determine_n_primal/dual and set_primal_offsets
 *     (v2rdm_solver.cc:3294-3846), BuildConstraints (v2rdm_solver.cc:1826-2074)
 *     and the cost-vector packing of the external-Hamiltonian constructor
 *     (v2rdm_solver.cc:164-193, tei.cc:139-199).
 *
 * All functions return V2H_OK (0) or a positive error code; no exception
 * crosses the boundary.  v2h_last_error() returns the message of the last
 * failure on the calling thread.  Host pointers are copied in and out; all
 * solver state lives in device memory owned by the handle.  A handle is not
 * re-entrant: calls on one handle must not overlap.
 */
#ifndef V2RDM_HIP_H
#define V2RDM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define V2H_OK 0
#define V2H_ERR_INVALID_ARGUMENT 1
#define V2H_ERR_HIP 2
#define V2H_ERR_OUT_OF_MEMORY 3
#define V2H_ERR_INTERNAL 4

/* Mirrors libsdp::SDPOptions as filled at v2rdm_solver.cc:426-431 from the
 * plugin options of plugin.cc:94-100, 217, 238-241. */
typedef struct v2h_sdp_options {
    double sdp_objective_convergence; /* E_CONVERGENCE, default 1e-6 */
    double sdp_error_convergence;     /* R_CONVERGENCE, default 1e-5 */
    double cg_convergence;            /* CG_CONVERGENCE, default 1e-9 */
    int cg_maxiter;                   /* CG_MAXITER, default 10000 */
    int maxiter;                      /* MAXITER: total outer iterations, default 50 */
    int mu_update_frequency;          /* MU_UPDATE_FREQUENCY, default 1000 */
    double mu0;                       /* initial penalty parameter, default 1 */
} v2h_sdp_options;

/* One fixed-orbital DQG problem (the v2RDMHelper external-Hamiltonian
 * constructor, python_helpers.cc:106 / v2rdm_solver.cc:147-195, extended with
 * per-orbital irrep labels; NULL irreps = C1 as in v2rdm_solver.cc:988). */
typedef struct v2h_problem_desc {
    int nmo;            /* active orbitals */
    int nalpha, nbeta;  /* active electrons */
    int multiplicity;   /* 2S+1; 0 = nalpha - nbeta + 1 (v2rdm_solver.cc:985-986) */
    int nirrep;         /* 1, 2, 4 or 8 */
    const int *orbital_irreps; /* [nmo] in the caller's orbital order, or NULL */
    int constrain_q2, constrain_g2;   /* POSITIVITY (plugin.cc:220, v2rdm_solver.cc:4081-4104) */
    int constrain_spin;               /* CONSTRAIN_SPIN (plugin.cc:306) */
    int constrain_sz;                 /* CONSTRAIN_SZ (plugin.cc:309) */
    int constrain_t1;                 /* POSITIVITY DQGT1 (v2rdm_solver.cc:4101-4104; needs q2 and g2):
                                         T1 blocks after G2, rows t1.cc:419-793 */
    int constrain_t2;                 /* POSITIVITY DQGT2 / DQGT1T2 (v2rdm_solver.cc:4105-4112):
                                         T2 blocks after T1, rows t2.cc:1236-1843 */
    int constrain_d3;                 /* CONSTRAIN_D3 (v2rdm_solver.cc:4122-4124): D3 blocks after
                                         T2, rows d3.cc:45-406 */
} v2h_problem_desc;

typedef struct v2h_status {
    double mu, energy_primal, energy_dual, primal_error, dual_error;
    int64_t oiter_total, iiter_total;  /* libsdp oiter_total() / iiter_total() (v2rdm_solver.cc:1546-1547) */
    int converged;                     /* libsdp is_converged() (v2rdm_solver.cc:1361) */
    int active, status_bits, last_cg_iterations;
} v2h_status;

/* Progress monitor, called once per outer iteration and instance; the fields
 * of libsdp's SDPProgressMonitorFunction (v2rdm_solver.cc:115-124) plus the
 * instance index. */
typedef void (*v2h_monitor_fn)(int instance, int oiter, int iiter, double energy_primal, double energy_dual,
                               double mu, double primal_error, double dual_error, void *data);

const char *v2h_last_error(void);
int v2h_default_options(v2h_sdp_options *opt);
int v2h_device_count(int *count);
/* Plain copy of n doubles between two fresh device buffers, repeated
 * `repeats` times with the library's copy kernel (k_copy); *ms_per_copy is the
 * mean time of one copy.  It moves 8 n bytes each way and calibrates the HBM
 * traffic counters (FETCH_SIZE, WRITE_SIZE) of a profile.  No counterpart in
 * the reference. */
int v2h_calibration_copy(int device, int64_t n, int repeats, double *ms_per_copy);

typedef struct v2h_batch v2h_batch;

/* A batch of independent SDPs on one GPU; count = 1 is the single-instance
 * replacement of new libsdp::BPSDPSolver(n_primal, n_dual, options). */
int v2h_batch_create(const v2h_problem_desc *descs, int count, const v2h_sdp_options *opt, int device,
                     v2h_batch **out);
int v2h_batch_destroy(v2h_batch *b);
int v2h_batch_count(const v2h_batch *b, int *count);
/* Replace the options.  A tighter E_CONVERGENCE or R_CONVERGENCE re-opens the
 * instances that had converged, so that the next v2h_batch_solve continues
 * them from their current x, y, z and mu to the new thresholds. */
int v2h_batch_set_options(v2h_batch *b, const v2h_sdp_options *opt);
/* Stopping rule of the CG solve inside each outer iteration: relative
 * residual ||r|| / ||rhs|| below factor x max(primal error, dual error) of the
 * previous outer iteration, capped at 1e-2 and at least CG_CONVERGENCE
 * (default factor 0.003, the oracle's).  libsdp's own schedule for this solve
 * (BPSDPSolver, not vendored; called at v2rdm_solver.cc:1340) is not known;
 * a larger factor solves each CG less tightly and changes the trajectory,
 * not the optimum.  factor in (0, 1]. */
int v2h_batch_set_cg_tolerance_factor(v2h_batch *b, double factor);

/* Sizes (determine_n_primal / determine_n_dual) and the PSD block list
 * (dimensions_, v2rdm_solver.cc:3729-3846) of one instance. */
int v2h_batch_sizes(const v2h_batch *b, int inst, int64_t *n_primal, int64_t *n_dual, int *n_blocks);
int v2h_batch_block_dims(const v2h_batch *b, int inst, int *dims);
/* Pitzer position -> caller orbital index (orbitals of one irrep contiguous). */
int v2h_batch_orbital_order(const v2h_batch *b, int inst, int *to_user);

/* Hamiltonian: h[nmo*nmo], g[nmo^4] chemists' (pq|rs) at ((p*n+q)*n+r)*n+s, in
 * the caller's orbital order; packed on the device into c. */
int v2h_batch_set_integrals(v2h_batch *b, int inst, const double *h, const double *g);
int v2h_batch_set_cost(v2h_batch *b, int inst, const double *c);
/* Density-fitted molecular Hamiltonian, contracted and packed on the device:
 * the DF branch of GetIntegrals / RepackIntegrals / FrozenCoreEnergy / TEI
 * (v2rdm_casscf/tei.cc:58-67, 139-275, 280-283), whose integrals are
 * TEI(i,j,k,l) = sum_Q qmo[Q*ntri + INDEX(i,j)] qmo[Q*ntri + INDEX(k,l)].
 *   nfull   orbitals of the integral space (nmo_ - nfrzv_ in the reference)
 *   h_full  [nfull*nfull] one-electron integrals, read on and above the
 *           diagonal (oei_full_sym_, tei.cc:124-132)
 *   qmo     [nq * ntri], ntri = nfull (nfull + 1) / 2: (Q|mn) in the layout
 *           of the reference's Qmo_ (INDEX(m,n) = packed lower triangle)
 *   core    [ncore] full-space indices of the doubly occupied (frozen and
 *           restricted) core orbitals; NULL when ncore = 0
 *   active  [nmo] full-space index of each active orbital, caller order
 *   efzc    out (may be NULL): the frozen-core energy, which the reference
 *           adds to c.x (v2rdm_solver.cc:1385-1386)
 * The one-electron part of c receives the core Fock terms of
 * FrozenCoreEnergy (tei.cc:229-268). */
int v2h_batch_set_df_integrals(v2h_batch *b, int inst, int nfull, const double *h_full, int nq, const double *qmo,
                               int ncore, const int *core, const int *active, double *efzc);
/* The binary integral and RDM files of the FCIDUMP option (v2RDMSolver::FCIDUMP,
 * v2rdm_casscf/fcidump.cc:60-243) for one instance, from density-fitted
 * integrals (the reference's FCIDUMP works with DF integrals only).  Each
 * file is a sequence of 24-byte records: a double and four int32 labels.
 *   int_path: every (pq|rs) with p <= q, r <= s, INDEX(p,q) <= INDEX(r,s) and
 *     equal pair symmetry, labels (p,q,r,s); then h_pq for p, q of one irrep,
 *     labels (p,q,0,0); then (enuc,0,0,0,0).  Labels are 1-based and order
 *     the orbitals core, active, virtual, each by irrep, then index.
 *   rdm_path: the spin-free 2-RDM D2ab(ij,kl) + D2ab(ji,lk) +
 *     s (D2aa + D2bb)(ij,kl) over the ab geminals (labels i,j,k,l relative
 *     to the core), then the spin-summed 1-RDM (labels i,j,0,0), of the
 *     instance's current primal vector.
 * full_irreps may be NULL (C1); core may be NULL when ncore = 0; active[u] is
 * the full-space index of the instance's orbital u.  The reference writes
 * after transforming to natural orbitals; pass integrals and x in the basis
 * wanted.  n_int / n_rdm (may be NULL) receive the record counts.  The pair
 * Gram matrix, (nfull (nfull + 1) / 2)^2 doubles, is formed on the device and
 * copied to the host once; beyond 64 GiB the call is refused. */
int v2h_batch_write_fcidump(v2h_batch *b, int inst, int nfull, const int *full_irreps, const double *h_full, int nq,
                            const double *qmo, int ncore, const int *core, const int *active, double enuc,
                            const char *int_path, const char *rdm_path, int64_t *n_int, int64_t *n_rdm);
/* One orbital-optimisation call of v2RDM-CASSCF on the instance's resident
 * density-fitted integrals (those of the last v2h_batch_set_df_integrals,
 * rotated by every call since): v2RDMSolver::RotateOrbitals
 * (v2rdm_casscf/v2rdm_solver.cc:3221-3292) with the DF optimiser
 * OrbitalOptimizer::optimize_orbitals (focas/orbopt_optimize.cc:25-110,
 * exact diagonal Hessian).  The current primal x gives the densities
 * (PackSpatialDensity, v2rdm_solver.cc:3110-3219); rotations between core,
 * active and virtual orbitals of one irrep (plus active-active ones when
 * opt->active_active) are optimised for up to opt->maxiter steps, the
 * integrals are rotated in place and c and E_fzc are repacked
 * (RepackIntegrals).  The SDP state (y, z, mu) is kept; unless the call
 * converged, the instance is marked unconverged.  full_irreps [nfull] (NULL: C1) label the full space; the
 * active orbitals must carry the instance's irreps.  res (may be NULL):
 * iterations, converged (the call-level test of v2rdm_solver.cc:3271-3274:
 * |energy change over the whole call| < e_convergence and the last gradient
 * norm < g_convergence; unless it holds the instance is marked unconverged;
 * the caller keeps it as a flag that is never reset, as orbopt_converged_ is),
 * rotation count,
 * gradient norm at the last evaluation, dE over the call, the energy c.x +
 * E_fzc afterwards and E_fzc. */
typedef struct v2h_orbopt_options {
    double g_convergence; /* ORBOPT_GRADIENT_CONVERGENCE, hilbert default 1e-4 */
    double e_convergence; /* ORBOPT_ENERGY_CONVERGENCE, default 1e-8 */
    int maxiter;          /* ORBOPT_MAXITER, default 10 */
    int active_active;    /* ORBOPT_ACTIVE_ACTIVE_ROTATIONS, default 0 */
} v2h_orbopt_options;
typedef struct v2h_orbopt_result {
    int iterations, converged, nrot;
    int evaluations, rotations; /* gradient evaluations; rotations applied (steps and undos) */
    double gnorm, dE, energy, efzc;
} v2h_orbopt_result;
int v2h_batch_optimize_orbitals(v2h_batch *b, int inst, const int *full_irreps, const v2h_orbopt_options *opt,
                                v2h_orbopt_result *res);
/* The orbital gradient dE/dkappa and the exact diagonal Hessian at the
 * current orbitals and primal x (ComputeG and ComputeH_diag of the DF
 * optimiser, focas/orbopt_gradient.cc:26-44, orbopt_hessian.cc:25-37), with
 * nothing rotated.  Rotation k turns orbitals p = pairs[2k] and q =
 * pairs[2k+1] into each other: K[p][q] = kappa, K[q][p] = -kappa, U = exp(K);
 * p is virtual and q active or core, or p active and q core, or (active-active)
 * p > q.  *nrot receives the pair count; the arrays are written only when
 * cap >= that count (call with cap = 0 to size them). */
int v2h_batch_orbital_gradient(v2h_batch *b, int inst, const int *full_irreps, int active_active, int32_t *pairs,
                               double *g, double *H, int cap, int *nrot);
/* The orbitals after optimisation: U [nfull*nfull] with new orbital r =
 * sum_s old_s U[s*nfull + r] relative to the orbitals of
 * v2h_batch_set_df_integrals (the reference multiplies its Ca by the same
 * rotation, orbopt_integraltransform_cpu.cc:35-58), and the rotated h_full
 * [nfull*nfull] and qmo [nq*ntri].  Any pointer may be NULL. */
int v2h_batch_get_df_orbitals(v2h_batch *b, int inst, double *U, double *h_full, double *qmo);
/* Primal guess; NULL = hilbert's TPDM_GUESS RANDOM (v2rdm_solver.cc:1799-1806). */
int v2h_batch_set_guess(v2h_batch *b, int inst, const double *x);
/* TPDM_GUESS HF (Guess, v2rdm_solver.cc:1711-1798; Q2_constraints_guess,
 * q2.cc:45-167; G2_constraints_guess, g2.cc:44-242): the determinant of the
 * first orbitals of the caller's order; nalpha orbitals of each spin are
 * occupied, as the external path's doccpi = nalpha (v2rdm_solver.cc:1004-1007)
 * makes the reference do; D200 left zero. */
int v2h_batch_set_guess_hf(v2h_batch *b, int inst);
/* y = z = 0, mu = mu0, iteration counters cleared. */
int v2h_batch_reset(v2h_batch *b, int inst);

/* Up to maxiter outer iterations for every unconverged instance (the
 * maxiter argument of SDPSolver::solve).  y, z and mu persist across calls. */
int v2h_batch_solve(v2h_batch *b, int maxiter, v2h_monitor_fn mon, void *data, int *n_converged);
int v2h_batch_status(v2h_batch *b, int inst, v2h_status *st);
double v2h_batch_last_solve_seconds(const v2h_batch *b);

/* Vectors of one instance, primal layout (v2rdm_solver.cc:3729-3846). */
int v2h_batch_get_x(v2h_batch *b, int inst, double *x);
int v2h_batch_get_y(v2h_batch *b, int inst, double *y);   /* libsdp get_y (checkpoint.cc:58-64) */
int v2h_batch_get_z(v2h_batch *b, int inst, double *z);   /* libsdp get_z */
int v2h_batch_set_y(v2h_batch *b, int inst, const double *y);  /* libsdp set_y (checkpoint.cc:106-120) */
int v2h_batch_set_z(v2h_batch *b, int inst, const double *z);  /* libsdp set_z */
int v2h_batch_get_mu(v2h_batch *b, int inst, double *mu);      /* libsdp get_mu */
int v2h_batch_set_mu(v2h_batch *b, int inst, double mu);       /* libsdp set_mu */
int v2h_batch_get_c(v2h_batch *b, int inst, double *c);
int v2h_batch_get_b(v2h_batch *b, int inst, double *bvec);

/* Reduced density matrices of the current primal x, orbital indices in the
 * caller's order (no frozen or restricted core).
 * get_opdm: dense n x n, type "A", "B" or "SUM" (v2RDMSolver::get_opdm,
 *   v2rdm_casscf/python_api.cc:374-394, which returns "SUM").
 * get_tpdm: dense spin-free 2-RDM, out[(i n + k) n^2 + j n + l]
 *   (v2RDMSolver::get_tpdm, python_api.cc:396-456; also for nirrep > 1).
 * get_*_sparse: the reference's entry lists (python_api.cc:39-371); index
 *   tuples into idx (2 or 4 per entry), values into val, at most cap entries
 *   written, *count = the full length (call with cap = 0 to size the arrays). */
int v2h_batch_get_opdm(v2h_batch *b, int inst, const char *type, double *out);
int v2h_batch_get_tpdm(v2h_batch *b, int inst, double *out);
int v2h_batch_get_opdm_sparse(v2h_batch *b, int inst, const char *type, int32_t *idx, double *val, int64_t cap,
                              int64_t *count);
int v2h_batch_get_tpdm_sparse(v2h_batch *b, int inst, const char *type, int32_t *idx, double *val, int64_t cap,
                              int64_t *count);

/* The constraint maps (bpsdp_Au / bpsdp_ATu, v2rdm_solver.cc:2995-3108). */
int v2h_batch_Au(v2h_batch *b, int inst, const double *u, double *out);
int v2h_batch_ATu(v2h_batch *b, int inst, const double *y, double *out);
/* v2h_batch_Au without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * input and output are staged in a 4 MiB scratch array that the library
 * keeps on each GPU until the process ends (its address is resolved on the
 * first call on a device): n_primal + n_dual of the instance must be at most
 * 524288 doubles, else V2H_ERR_INVALID_ARGUMENT.  Every call on one GPU shares
 * that scratch: calls that overlap on one GPU (from several threads or
 * handles) race and must be serialised by the caller.  Same result as
 * v2h_batch_Au, bit for bit. */
int v2h_batch_Au_noalloc(v2h_batch *b, int inst, const double *u, double *out);

/* HIP-event timing of the kernels launched by v2h_batch_solve, recorded on
 * the library's stream.  ms[16] and counts[16] per kernel kind: 0 iteration
 * start (A [c - z, x, A^T y]), 1 CG A^T p, 2 CG A q, 3 A^T y and U, 4..10 PSD
 * projection classes: Jacobi for blocks up to 16, tridiagonal for blocks up to
 * 64 / 128 / 192 / 224 / 256, Jacobi beyond 256 (matrix in global memory);
 * 11 the device part of v2h_batch_set_df_integrals (from the first
 * contraction kernel to the packed c), 12 the pair Gram matrix of
 * v2h_batch_write_fcidump, 13 the device work of v2h_batch_optimize_orbitals
 * besides the intake (densities, Fock and Hessian products, exponential,
 * integral transform), 14 the warm-started refinement of the blocks of
 * 65..256 rows (all its launches of one outer iteration), 15 the eigenvector
 * formation of the blocks that took the tridiagonal fallback (with the
 * refinement on, kinds 6..9 time only the fallback launches).  Cumulative
 * since the last reset.  full_ms / full_counts (may be NULL) cover only the launches
 * made while every instance was running (for CG kernels: still iterating)
 * at the previous host poll, i.e. launches that processed the whole batch. */
int v2h_batch_set_profiling(v2h_batch *b, int on);
/* hipGraph replay inside v2h_batch_solve (the graphs field of v2h_tuning):
 * mode 1 captures the first CG chunk of every outer iteration (its k_cg_at /
 * k_cg_a / k_cg_reduce launches) once per set of running instances and step
 * count and replays it with one hipGraphLaunch (with the kernel timers on,
 * v2h_batch_set_profiling, the launches go one by one); the results are the
 * bits of mode 0 (default, every kernel launched one by one).  Mode 2 also
 * captures k_post and the whole projection (forked onto the side streams and
 * joined back), with the eigenvector formation as a branch of the next CG
 * graph; the same bits again.
 * The libsdp loop it replaces launches nothing (v2rdm_solver.cc:1305-1361). */
int v2h_batch_set_graphs(v2h_batch *b, int mode);

/* Launch configuration of v2h_batch_solve.  None of it changes a result bit:
 * the same kernels run on the same data in the same dependency order, and
 * the extra CG steps of a larger first chunk are skipped by the instances
 * whose CG has stopped.  hilbert_amd/autotune.py times candidates on a batch
 * and keeps the fastest per batch signature.  No counterpart in the
 * reference, whose libsdp loop runs on the host. */
typedef struct {
    int class_lane[7];    /* stream of each PSD size class when not refined:
                           * Jacobi <= 16, tridiagonal <= 64 / 128 / 192 /
                           * 224 / 256, Jacobi > 256; 0 = the batch's main
                           * stream, 1..3 its side streams */
    int fallback_lane[7]; /* stream of the tridiagonal fallback of the refined
                           * classes (entries 2..5 used) */
    int refine_lane;      /* side stream (1..3) of the refinement and the
                           * eigenvector formation */
    int cg_margin;        /* first CG chunk of an outer iteration: the largest
                           * CG step count of the previous one + cg_margin
                           * (0..64); then chunks of 4 */
    int xcd_remap;        /* 1: gather workgroups dealt to the 8 XCDs in
                           * contiguous ranges (one instance per L2); 0
                           * (default): in launch order */
    int graphs;           /* hipGraph replay: 0 off, 1 first CG chunk, 2
                           * also k_post and the projection
                           * (v2h_batch_set_graphs) */
    int pipeline;         /* > 0: while at least this many instances run,
                           * two groups of them take turns on the main stream
                           * (pre, CG, post) and each group's projection runs
                           * on the side streams beside the other group's CG;
                           * 0: lock-step (every instance together); ignored
                           * with graphs */
} v2h_tuning;
int v2h_batch_get_tuning(const v2h_batch *b, v2h_tuning *t);
int v2h_batch_set_tuning(v2h_batch *b, const v2h_tuning *t);
int v2h_batch_kernel_times(v2h_batch *b, double *ms, int64_t *counts, double *full_ms, int64_t *full_counts);
int v2h_batch_reset_kernel_times(v2h_batch *b);
/* Launch statistics: [kernel launches so far, A gather workgroups,
 * A^T gather workgroups, projection blocks] */
int v2h_batch_launch_stats(v2h_batch *b, int64_t *stats);

/* How the PSD blocks of 65..256 rows are projected (the eigendecomposition
 * libsdp applies to every block of dimensions_, v2rdm_solver.cc:1340, block
 * list v2rdm_solver.cc:3729-3846):
 *   V2H_PROJECT_REFINED (default): the eigenvectors of the previous outer
 *     iteration are refined by FP64 matrix-core products (iterative
 *     refinement of Ogita and Aishima, at most five steps by default,
 *     accepted when the last correction is below 2e-7); blocks without
 *     eigenvectors yet, or
 *     whose refinement is not accepted, take the tridiagonal eigensolver,
 *     whose eigenvectors seed the next iteration;
 *   V2H_PROJECT_TRIDIAGONAL: every block takes the tridiagonal eigensolver.
 * Smaller and larger blocks are not affected. */
#define V2H_PROJECT_TRIDIAGONAL 0
#define V2H_PROJECT_REFINED 1
int v2h_batch_set_projection(v2h_batch *b, int mode);
/* Over the blocks of 65..256 rows since the batch was created: [block
 * projections done by the refinement, by the tridiagonal fallback, refinements
 * rejected for a diverging or non-finite correction, for a correction still
 * above accept_e after max_steps steps, for clusters left coupled, FP64 flops
 * of the refinement's d x d x d products (2 d^3 each)]; stats[6]. */
int v2h_batch_projection_stats(v2h_batch *b, int64_t *stats);
/* Thresholds of the refinement: eigenvalue estimates closer than cluster_rel
 * x max|l| form a cluster (default 1e-8); a step is accepted when its largest
 * correction is at most accept_e (2e-7) and the couplings left inside
 * clusters at most accept_cpl x max|l| (1e-13); a block not accepted after
 * max_steps steps (1..8, default 5) takes the tridiagonal fallback. */
int v2h_batch_set_refine_params(v2h_batch *b, double cluster_rel, double accept_e, double accept_cpl,
                                int max_steps);

/* Adjoint of the cost packing: dE/dh[nmo*nmo], dE/dg[nmo^4] (caller order)
 * for E = c . xbar; xbar = NULL uses the current primal solution x. */
int v2h_batch_cost_adjoint(v2h_batch *b, int inst, const double *xbar, double *dh, double *dg);

/* hilbert's RDM files for downstream codes, from instance inst's current x:
 * the PSIO records of WriteActiveTPDM / WriteTPDM / WriteTPDMSpinFree /
 * WriteOPDM (write_tpdm.cc:521-595, 40-300, 312-519, write_opdm.cc:40-119)
 * and the IWL TPDM of WriteTPDM_IWL (write_tpdm_iwl.cc:45-200), chosen by
 * `which` (V2H_RDM_* flags, the options that make hilbert call each writer,
 * v2rdm_solver.cc:1474-1509) and written in hilbert's order.  Orbital labels
 * are full-space indices: active[u] for active orbital u (caller order),
 * core[] the doubly occupied core orbitals in hilbert's order (by irrep, then
 * index).  Each file is a sequence of PSIO-style keyed entries {char key[80],
 * int64 nbytes, bytes}:
 *   <prefix>.d2aa / .d2bb / .d2ab: "D2aa" / "D2bb" / "D2ab" records {int i, j,
 *     k, l; double value} = D2(ij, kl), "length" (int64); TPDM_FULL adds the
 *     core contributions and "NUMBER ACTIVE ORBITALS", "ACTIVE ORBITALS",
 *     "NUMBER INACTIVE ORBITALS", "INACTIVE ORBITALS" (with both flags the
 *     full records replace the active ones, as in hilbert);
 *   <prefix>.d2: "D2" spin-free records, "length" and the orbital lists
 *     (hilbert's same-spin sign: -1 as soon as i > j or k > l);
 *   <prefix>.d1a / .d1b: "D1a" / "D1b" records {int i, j; double value},
 *     "length";
 *   <prefix>.aa.iwl / .bb.iwl / .ab.iwl: "IWL Buffers", psi4 libiwl buffers
 *     {int lastbuf, int inbuf, int16 labels[4 x 2980], double values[2980]},
 *     values above 1e-14 in magnitude, labels (p, q, r, s) in chemists'
 *     order, the aa / bb values x 1/4 as hilbert writes them.
 * counts[9]: D1a, D1b, D2aa, D2bb, D2ab, spin-free D2 records and IWL aa, bb,
 * ab values; -1 for a file not written. */
#define V2H_RDM_TPDM_ACTIVE 1     /* TPDM_WRITE */
#define V2H_RDM_TPDM_FULL 2       /* TPDM_WRITE_FULL */
#define V2H_RDM_TPDM_SPIN_FREE 4  /* TPDM_WRITE_SPIN_FREE */
#define V2H_RDM_OPDM_FULL 8       /* OPDM_WRITE_FULL */
#define V2H_RDM_TPDM_IWL 16       /* DERTYPE FIRST */
int v2h_batch_write_rdm_files(v2h_batch *b, int inst, int ncore, const int *core, const int *active,
                              const char *prefix, int which, int64_t *counts);

/* ---- BPSDP with caller-supplied maps ----
 * The generic libsdp contract: BPSDPSolver(n_primal, n_dual, options)
 * (v2rdm_solver.cc:433) and solve(x, b, c, block_dims, maxiter, Au, ATu,
 * monitor, print_level, data) (v2rdm_solver.cc:1340), with the callback types
 * of v2rdm_solver.cc:97-124: Au writes out[n_dual] = A in, ATu writes
 * out[n_primal] = A^T in, both on host memory.  For solvers whose maps are not
 * the built-in DQG ones (v2RDM-DOCI, other constraint families).  The CG
 * vectors live on the host beside the maps; the PSD projection of the blocks
 * (sizes in block_dims, row-major and contiguous in x) runs on the GPU.
 * y, z and mu persist across solve calls, as libsdp's do. */
typedef void (*v2h_map_fn)(double *out, double *in, void *data);
/* libsdp's SDPProgressMonitorFunction (v2rdm_solver.cc:115-124), called after
 * every outer iteration with the print_level given to v2h_sdp_solve, so that
 * hilbert's bpsdp_monitor (which prints only for print_level > 0) passes
 * through unchanged; iiter = CG steps of that outer iteration. */
typedef void (*v2h_sdp_monitor_fn)(int print_level, int oiter, int iiter, double energy_primal, double energy_dual,
                                   double mu, double primal_error, double dual_error, void *data);
typedef struct v2h_sdp v2h_sdp;
int v2h_sdp_create(int64_t n_primal, int64_t n_dual, const v2h_sdp_options *opt, int device, v2h_sdp **out);
int v2h_sdp_destroy(v2h_sdp *s);
int v2h_sdp_set_options(v2h_sdp *s, const v2h_sdp_options *opt);
/* Up to maxiter outer iterations from x, which receives the result;
 * *converged = libsdp is_converged() (v2rdm_solver.cc:1361). */
int v2h_sdp_solve(v2h_sdp *s, double *x, const double *b, const double *c, const int *block_dims, int n_blocks,
                  int maxiter, v2h_map_fn Au, v2h_map_fn ATu, v2h_sdp_monitor_fn mon, int print_level, void *data,
                  int *converged);
int v2h_sdp_status(const v2h_sdp *s, v2h_status *st);
int v2h_sdp_get_y(const v2h_sdp *s, double *y);   /* libsdp get_y (checkpoint.cc:58-64) */
int v2h_sdp_get_z(const v2h_sdp *s, double *z);   /* libsdp get_z */
int v2h_sdp_set_y(v2h_sdp *s, const double *y);   /* libsdp set_y (checkpoint.cc:106-120) */
int v2h_sdp_set_z(v2h_sdp *s, const double *z);   /* libsdp set_z */
int v2h_sdp_get_mu(const v2h_sdp *s, double *mu); /* libsdp get_mu */
int v2h_sdp_set_mu(v2h_sdp *s, double mu);        /* libsdp set_mu */

/* ---- host-only inspection of one problem structure (no GPU needed) ----
 * The tables the device path is built from: geminal tables
 * (build_mapping_arrays, basis.cc:427-533), the constraint matrix A with b in
 * CSR (the rows of bpsdp_Au, v2rdm_solver.cc:2995-3050, and BuildConstraints,
 * v2rdm_solver.cc:1826-2074), and the cost-vector gather map
 * (v2rdm_solver.cc:164-193).  Orbital indices are Pitzer positions. */
typedef struct v2h_structure v2h_structure;
int v2h_structure_create(const v2h_problem_desc *desc, v2h_structure **out);
int v2h_structure_destroy(v2h_structure *s);
int v2h_structure_sizes(const v2h_structure *s, int64_t *n_primal, int64_t *n_dual, int64_t *nnz, int *n_blocks);
int v2h_structure_block_dims(const v2h_structure *s, int *dims);
/* gems_ab[nirrep], gems_aa[nirrep], bas_ab[2 sum gems_ab], bas_aa[2 sum gems_aa] */
int v2h_structure_tables(const v2h_structure *s, int *gems_ab, int *gems_aa, int *bas_ab, int *bas_aa);
int v2h_structure_constraints(const v2h_structure *s, int64_t *row_ptr, int32_t *col, double *val, double *b);
int v2h_structure_cost_map(const v2h_structure *s, int32_t *i1, int32_t *i2);
/* SELL-64 statistics: [A slices, A entries, A long rows, A long entries,
 * A^T slices, A^T entries, A^T long rows, A^T long entries] */
int v2h_structure_gather_stats(const v2h_structure *s, int64_t *stats);
/* RDM exports of a given primal vector x (as v2h_batch_get_opdm / _tpdm). */
int v2h_structure_opdm(const v2h_structure *s, const double *x, const char *type, double *out);
int v2h_structure_tpdm(const v2h_structure *s, const double *x, double *out);
/* The TPDM_GUESS HF primal vector (as v2h_batch_set_guess_hf). */
int v2h_structure_hf_guess(const v2h_structure *s, double *x);
/* v2h_batch_write_rdm_files for a primal vector x of this structure (host only). */
int v2h_structure_write_rdm_files(const v2h_structure *s, const double *x, int ncore, const int *core,
                                  const int *active, const char *prefix, int which, int64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
