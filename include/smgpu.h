/*
 * smgpu — C ABI of the MI355X M-scheme shell-model eigensolver.
 *
 * The boundary replaces the body of NuclearToolkit.jl's `main_sm` path from the
 * basis enumeration through the Lanczos solve (reference paths relative to its
 * repository root).  Plain C types only; every call returns an int32 status
 * (SMGPU_OK or a negative code) and never throws; smgpu_last_error() gives the
 * text of the last failure on the calling thread.  Handles own their host and
 * device memory; callers own every pointer they pass.  Indices cross 0-based.
 * Vectors use the reference's global M-scheme order:
 *   index = tdims[b] + ip * nN[b] + in
 * (block b by ascending proton M; proton and neutron determinants of a block in
 * descending bit-integer order, first m-state = most significant bit).
 */
#ifndef SMGPU_H
#define SMGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMGPU_OK 0
#define SMGPU_EINVAL (-1)   /* bad argument or malformed interaction file */
#define SMGPU_EIO (-2)      /* file not readable */
#define SMGPU_EDEVICE (-3)  /* HIP runtime / device failure */
#define SMGPU_ENOMEM (-4)   /* does not fit (host or HBM) */
#define SMGPU_ENOTCONV (-5) /* iteration limit reached */
#define SMGPU_ECANCELED (-6)

#define SMGPU_FP64 0        /* FP64 H·v and Krylov basis: the plain FP64 TRL */
#define SMGPU_MIXED 1       /* FP32 H·v and basis, FP64 recurrence/reorth/eigensolves */

typedef struct smgpu_problem_s* smgpu_problem;  /* interaction + basis + host tables */
typedef struct smgpu_ham_s* smgpu_ham;          /* device-resident tables on one GPU */
typedef struct smgpu_server_s* smgpu_server;    /* persistent serving loop */
typedef struct smgpu_moments_s* smgpu_moments;  /* M1 / E2 one-body operators on a basis */
typedef struct smgpu_transitions_s* smgpu_transitions;  /* M1 / E2 operators between two bases */

typedef struct {
    int32_t Z, N, A, vp, vn, twoM;
    int32_t lp, ln, cp, cn;          /* orbits and core of the interaction file */
    int32_t nblocks;
    int32_t n_msps_p, n_msps_n;      /* m-states per species */
    int64_t mdim;
    int64_t n_psd, n_nsd;            /* proton / neutron determinants over all blocks */
    int64_t n_pp_ops, n_nn_ops, n_pn_ops;   /* two-body operator lists (prep_Hamil_T1 / _pn) */
    int64_t n_hop_entries;           /* one-body hop-table entries (both species) */
} smgpu_problem_info;

typedef struct {
    int32_t num_ev;      /* lowest eigenpairs wanted (main_sm num_ev), 1..64 */
    int32_t lm;          /* Krylov vectors kept (main_sm lm, default 300), up to 1000 */
    int32_t ls;          /* vectors kept at a thick restart (main_sm ls, default 20), 1..64, < lm - 1 */
    int32_t num_history; /* Ritz-value history (main_sm num_history, default 3) */
    double tol;          /* convergence: all(en_prev - en_new < tol) (default 1e-8) */
    int32_t precision;   /* SMGPU_FP64 or SMGPU_MIXED */
    int32_t refine;      /* MIXED only: FP64 Rayleigh quotient of the Ritz vectors */
    int32_t max_iter;    /* safety limit on Lanczos iterations (0 = 100000) */
    uint64_t seed;       /* start vector seed */
    int32_t reorth;      /* SMGPU_REORTH_*: re-orthogonalisation of the TRL basis */
    int32_t steps;       /* SMGPU_STEPS_*: how the Lanczos steps of a solve reach the device */
} smgpu_solve_params;

/* SMGPU_STEPS_GRAPH: the kernels of one Lanczos step (H·v, reorthogonalisation,
 * Ritz values and convergence test) are captured once per solve as a hipGraph
 * and replayed every step, reading the iteration from device memory; LAUNCH: each
 * kernel is launched; AUTO: the library's choice by basis size.  Results are
 * bit-identical.  (Single solves; lockstep batches and J-projected solves launch.) */
#define SMGPU_STEPS_AUTO 0
#define SMGPU_STEPS_LAUNCH 1
#define SMGPU_STEPS_GRAPH 2

/* Re-orthogonalisation of each new Lanczos vector against the stored basis:
 * DEFAULT is FULL for FP64 (the reference's ReORTH at every step) and PARTIAL for
 * MIXED.  PARTIAL follows the loss of orthogonality with Simon's recurrence for
 * omega_jk = v_j . v_k and re-orthogonalises a step (and the next one) only when
 * max |omega| exceeds sqrt(eps) of the basis precision, keeping the basis
 * semi-orthogonal (H.D. Simon, Math. Comp. 42 (1984) 115; Lin. Alg. Appl. 61 (1984) 101). */
#define SMGPU_REORTH_DEFAULT 0
#define SMGPU_REORTH_FULL 1
#define SMGPU_REORTH_PARTIAL 2

typedef struct {
    int32_t iterations;  /* elit of the last Lanczos sweep */
    int32_t n_hv;        /* H·v applications in the working precision */
    int32_t n_hv64;      /* FP64 H·v used by the refinement */
    int32_t n_restart;
    double wall_ms;
    double ritz[64];     /* en[1]: Ritz values of the last convergence check */
    int32_t n_reorth2;   /* iterations that needed a second Gram-Schmidt pass */
    int32_t n_jproj;     /* J-projected solves: vectors that were projected */
    int32_t n_reorth;    /* iterations that ran a re-orthogonalisation pass */
    int32_t n_fallback;  /* 1: a mixed solve whose Ritz values disagreed with the FP64
                            Rayleigh quotients (lost orthogonality) was redone with
                            reorthogonalisation at every step */
} smgpu_solve_stats;

/* library */
int32_t smgpu_version(void);
const char* smgpu_last_error(void);
int32_t smgpu_device_count(int32_t* n);

/* readsmsnt + getZNA + construct_msps + occ + prep_Hamil_T1/prep_2bjumps +
 * prep_Hamil_pn/prep_pn (shellmodel_main.jl:96-147, 293-365, 373-418, 580-730,
 * 769-848, 1115-1199).  twoM < 0 selects the main_sm default A%2 (:110).
 * ignore_massop mirrors readsmsnt's keyword (:293).  nthreads <= 0: all cores. */
int32_t smgpu_problem_create(const char* snt_path, const char* nucleus, int32_t twoM, int32_t ignore_massop,
                             int32_t nthreads, smgpu_problem* out);
int32_t smgpu_problem_destroy(smgpu_problem p);
int32_t smgpu_problem_get_info(smgpu_problem p, smgpu_problem_info* info);
/* Replace the interaction of an existing problem (same model space, basis and
 * index tables) — the shared-basis loop of prepEC (eigenvector_continuation.jl:139-212). */
int32_t smgpu_problem_set_interaction(smgpu_problem p, const char* snt_path);

/* Basis export for bit-exact checks against occ (shellmodel_main.jl:769-848):
 * tdims[nblocks+1], Mp[nblocks], Mn[nblocks], pbits/nbits concatenated per
 * block (n_psd / n_nsd entries, block order), psd_off/nsd_off[nblocks+1]. */
int32_t smgpu_basis_export(smgpu_problem p, int64_t* tdims, int32_t* Mp, int32_t* Mn, int64_t* pbits,
                           int64_t* psd_off, int64_t* nbits, int64_t* nsd_off);

/* One-body hop tables in a canonical form for comparison with calc_1b_jumps!
 * (shellmodel_main.jl:1169-1199): every (species, input block, output block,
 * input SD, output SD, creation m-state, annihilation m-state, phase bit of the
 * reference's calc_phase_1!).  Pass NULL arrays to get the count in *n. */
int32_t smgpu_hops_export(smgpu_problem p, int64_t* n, int32_t* species, int32_t* bi, int32_t* bf, int32_t* i,
                          int32_t* f, int32_t* a, int32_t* c, int8_t* phase);

/* Like-particle matrix entries per block (the symmetric matrix that
 * prep_2bjumps' entries define, shellmodel_main.jl:1115-1137): species 0 = pp,
 * 1 = nn; rows (block, f, i, value).  NULL arrays: count in *n. */
int32_t smgpu_pair_export(smgpu_problem p, int32_t species, int64_t* n, int32_t* blk, int32_t* f, int32_t* i,
                          double* val);

/* Upload tables to a device (the "uploaded once" step). */
int32_t smgpu_ham_create(smgpu_problem p, int32_t device, smgpu_ham* out);
int32_t smgpu_ham_destroy(smgpu_ham h);
/* Device tables of J^2 on the problem's basis (the operator of prep_J /
 * operate_J!, shellmodel_main.jl:149-157, 226-230, built as the interaction
 * sum_a j_a(j_a+1) n_a + sum_{ab,J} [J(J+1) - ja(ja+1) - jb(jb+1)] A+_J(ab) A_J(ab)):
 * smgpu_apply_H on this handle applies J^2, so <v|J^2|v> = J(J+1) of main_sm's Js. */
int32_t smgpu_ham_create_j2(smgpu_problem p, int32_t device, smgpu_ham* out);
/* Re-upload the coefficients after smgpu_problem_set_interaction (index tables stay). */
int32_t smgpu_ham_update_interaction(smgpu_ham h, smgpu_problem p);
/* A handle for another interaction on base's basis: it shares base's device index
 * tables and uploads only p's coefficients (the per-sample interaction tables of
 * prepEC on one basis, eigenvector_continuation.jl:140-175).  p must have base's
 * basis and pp / nn sparsity pattern.  The index lives while any handle holds it. */
int32_t smgpu_ham_create_variant(smgpu_ham base, smgpu_problem p, smgpu_ham* out);

/* Bytes of every array the H·v kernel streams per call (roofline accounting).
 * names: newline-separated names written into names_buf (may be NULL). */
int32_t smgpu_table_bytes(smgpu_ham h, int32_t precision, int32_t max_arrays, int64_t* bytes, int32_t* n_arrays,
                          char* names_buf, int32_t names_len);
/* LDS bytes one one-vector H·v launch of `precision` moves, by part: parts[0] own
 * rows staged and pn sums moved to natural order, [1] diagonal reads, [2] nn reads, [3] pn staging writes, [4] pn
 * walk reads.  With the launch time this gives the kernel's LDS roofline.  No
 * counterpart in the reference (measurement only). */
int32_t smgpu_lds_bytes(smgpu_ham h, int32_t precision, int64_t* parts);
/* FMAs per H·v: [pn terms incl. padding, pn terms exact, pp, nn, diagonal] */
int32_t smgpu_work_terms(smgpu_ham h, int64_t* terms5);

/* operate_H! (lanczos_methods.jl:153-248), as w = H v (the reference accumulates
 * into a zeroed twf).  Device pointers (precision FP64: double, MIXED: float),
 * on the handle's stream unless `stream` is given (hipStream_t). */
int32_t smgpu_apply_H(smgpu_ham h, int32_t precision, const void* v_dev, void* w_dev, void* stream);
/* Host-buffer convenience: FP64 in/out, computed in `precision` on the device. */
int32_t smgpu_apply_H_host(smgpu_ham h, int32_t precision, const double* v, double* w);
/* H·V on nvec vectors at once (bl_operate_H!, lanczos_methods.jl:435-548): w_b = H v_b
 * for v_b = v_dev + b*ldv, w_b = w_dev + b*ldw (elements; ld >= mdim when nvec > 1).
 * Launches of up to four vectors share every table load where the vectors' LDS fits
 * the workgroup's occupancy (up to four on narrow neutron blocks, two on wide FP32
 * ones, else one per launch); each vector's result is bit-identical to smgpu_apply_H
 * on it. */
int32_t smgpu_apply_H_multi(smgpu_ham h, int32_t precision, int32_t nvec, const void* v_dev, int64_t ldv,
                            void* w_dev, int64_t ldw, void* stream);
/* Host-buffer convenience: v and w are nvec x mdim row-major FP64. */
int32_t smgpu_apply_H_host_multi(smgpu_ham h, int32_t precision, int32_t nvec, const double* v, double* w);
/* Timing of smgpu_apply_H_multi on resident vectors: milliseconds per call (all nvec). */
int32_t smgpu_time_apply_H_multi(smgpu_ham h, int32_t precision, int32_t nvec, int32_t reps, double* ms);
/* w_b = H_b v_b for handles that share one index (smgpu_ham_create_variant; a handle
 * may repeat): vectors of one handle share launches as in smgpu_apply_H_multi,
 * vectors of different handles are paired, where their LDS fits, into launches that
 * share the index loads and entry decodes with per-vector coefficients.  Each result is bit-identical to
 * smgpu_apply_H(hs[b], ...) on v_b.  Device pointers in the launch precision; no
 * counterpart in the reference (one interaction per operate_H! call). */
int32_t smgpu_apply_H_variants(const smgpu_ham* hs, int32_t precision, int32_t nvec, const void* const* v_dev,
                               void* const* w_dev, void* stream);
/* Host-buffer convenience: v and w are nvec x mdim row-major FP64. */
int32_t smgpu_apply_H_host_variants(const smgpu_ham* hs, int32_t precision, int32_t nvec, const double* v,
                                    double* w);
/* Timing of smgpu_apply_H_variants on resident vectors: milliseconds per call (all nvec). */
int32_t smgpu_time_apply_H_variants(const smgpu_ham* hs, int32_t precision, int32_t nvec, int32_t reps, double* ms);
/* Timing: `reps` launches of H·v on resident device vectors, HIP events on the
 * launch stream; average milliseconds per launch in *ms. */
int32_t smgpu_time_apply_H(smgpu_ham h, int32_t precision, int32_t reps, double* ms);

/* TRL + Ritz assembly (lanczos_methods.jl:1-91, 250-307; shellmodel_main.jl:199-224):
 * evals[num_ev] gets main_sm's en[1]; evecs (num_ev * mdim, may be NULL) the
 * normalised Ritz vectors. */
void smgpu_default_params(smgpu_solve_params* p);
int32_t smgpu_trl_solve(smgpu_ham h, const smgpu_solve_params* params, double* evals, double* evecs,
                        smgpu_solve_stats* stats);

/* Same as smgpu_trl_solve with a caller-supplied start vector v0[mdim] (main_sm's
 * in_wf start, shellmodel_main.jl:194-198); v0 == NULL uses params->seed. */
int32_t smgpu_trl_solve_v0(smgpu_ham h, const smgpu_solve_params* params, const double* v0, double* evals,
                           double* evecs, smgpu_solve_stats* stats);
/* J-projected TRL: main_sm with target_J = [twoJ] (shellmodel_main.jl:111-114, TRL's
 * doubleLanczos with Jlanczos, lanczos_methods.jl:1-91, 114-151).  h and j2 (from
 * smgpu_ham_create_j2) must be built on a problem with 2M = twoJ; the lowest num_ev
 * states with J = twoJ/2.  Every new Lanczos vector whose <J^2> is off J(J+1) by more
 * than 1e-7 (Jlanczos' Jtol) is replaced by its exact projection on that J. */
int32_t smgpu_trl_solve_j(smgpu_ham h, smgpu_ham j2, int32_t twoJ, const smgpu_solve_params* params,
                          const double* v0, double* evals, double* evecs, smgpu_solve_stats* stats);
/* nb independent TRL solves on one Hamiltonian in lockstep (no counterpart in the
 * reference, which solves one start vector per main_sm call): every round issues
 * the H·v of all unfinished members in shared launches (smgpu_apply_H_multi's
 * kernels), then each member runs the rest of its Lanczos step on its own stream.
 * params, v0 (each NULL or mdim doubles; v0 itself may be NULL), evals (num_ev
 * doubles each), evecs (each NULL or num_ev*mdim; evecs itself may be NULL), stats
 * and status (may be NULL) are arrays of nb.  All members share one precision.
 * Each member's results are bit-identical to smgpu_trl_solve_v0 with its params.
 * Returns the first failing member's code (status[] holds every member's). */
int32_t smgpu_trl_solve_batch(smgpu_ham h, int32_t nb, const smgpu_solve_params* params, const double* const* v0,
                              double* const* evals, double* const* evecs, smgpu_solve_stats* stats,
                              int32_t* status);
/* smgpu_trl_solve_batch over interaction variants on one basis: member b solves
 * hs[b] (handles sharing one index, smgpu_ham_create_variant; a handle may repeat),
 * with each round's H·v in smgpu_apply_H_variants' launches (prepEC's samples,
 * eigenvector_continuation.jl:140-175, solved together instead of one after
 * another).  Each member's results are bit-identical to smgpu_trl_solve_v0(hs[b]). */
int32_t smgpu_trl_solve_variants(const smgpu_ham* hs, int32_t nb, const smgpu_solve_params* params,
                                 const double* const* v0, double* const* evals, double* const* evecs,
                                 smgpu_solve_stats* stats, int32_t* status);
/* Thick-restart block Lanczos: main_sm with is_block=true and block size q
 * (TRBL, lanczos_methods.jl:312-420, with bl_operate_H!, bl_ReORTH, bl_QR! and
 * bl_ThickRestart :422-617; shellmodel_main.jl:167-191, 211-224).  params->lm is the
 * basis cap (lm/q blocks), params->ls the kept Ritz vectors (q*(ls/q) <= 32).
 * v0: q start vectors (q*mdim, read_appwav's block start) or NULL for seeded
 * random ones.  j2 (from smgpu_ham_create_j2) may be NULL; with it, every new
 * block vector off J(J+1) by more than 1e-7 is projected on J = twoJ/2 (the
 * doubleLanczos filter of TRBL) and the lowest num_ev states of that J result.
 * If the block Krylov space is exhausted (it spans the basis or an invariant
 * subspace) the solve stops with T the exact projection on it. */
int32_t smgpu_trbl_solve(smgpu_ham h, smgpu_ham j2, int32_t twoJ, int32_t q, const smgpu_solve_params* params,
                         const double* v0, double* evals, double* evecs, smgpu_solve_stats* stats);
/* nsteps Lanczos steps of the TRL recurrence from v0 (no convergence test or
 * restart): alpha[nsteps], beta[nsteps] of the projected tridiagonal matrix. */
int32_t smgpu_lanczos_steps(smgpu_ham h, int32_t precision, int32_t nsteps, const double* v0, double* alpha,
                            double* beta);
/* smgpu_lanczos_steps without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same inputs,
 * outputs and results as smgpu_lanczos_steps (the same kernels in the same
 * order); every device buffer of the call (Krylov basis, work vectors, partial
 * sums, flags) lives in one static device array of smgpu_lanczos_scratch_bytes()
 * bytes (64 MiB) that each device holds until the process ends, and everything
 * runs on the handle's stream.  A size whose buffers do not fit returns
 * SMGPU_EINVAL (about (nsteps + 8) * mdim * 8 bytes in FP64; mixed precision
 * stores the basis in FP32).  The array keeps the last call's contents; calls on one
 * device share it and are serialized by a host mutex inside the library (calls
 * from separate processes have separate arrays). */
int32_t smgpu_lanczos_steps_noalloc(smgpu_ham h, int32_t precision, int32_t nsteps, const double* v0,
                                    double* alpha, double* beta);
int64_t smgpu_lanczos_scratch_bytes(void);

/* ------------------------------------------------------------------ moments
 * main_sm's calc_moment (shellmodel_main.jl:243-246): the magnetic dipole and
 * electric quadrupole moments of Ritz vectors.
 * smgpu_moments_create builds prep_obtr_diag's M = 0 one-body operators
 * (transit.jl:82-245) on the problem's basis and uploads them to `device`:
 * gfactors[4] = (g_l proton, g_l neutron, g_s proton, g_s neutron) and
 * effcharge[2] = (e_p, e_n) as main_sm's keywords (defaults 1, 0, 5.586, -3.826
 * and 1.5, 0.5); bpar <= 0 takes main_sm's oscillator length (init_ho_by_mass,
 * transit.jl:49-60, formula 2 for 16 <= A <= 40, shellmodel_main.jl:104-107). */
int32_t smgpu_moments_create(smgpu_problem p, int32_t device, const double* gfactors, const double* effcharge,
                             double bpar, smgpu_moments* out);
int32_t smgpu_moments_destroy(smgpu_moments m);
/* w = O v with O the M1 (which = 0) or E2 (1) operator (op_obtr_diag,
 * transit.jl:247-283), FP64 device vectors in the M-scheme order, on `stream`
 * (NULL: the handle's own) */
int32_t smgpu_moments_apply(smgpu_moments m, int32_t which, const double* v_dev, double* w_dev, void* stream);
/* the same on host vectors: nvec vectors of mdim amplitudes each */
int32_t smgpu_moments_apply_host(smgpu_moments m, int32_t which, int32_t nvec, const double* v, double* w);
/* eval_moment (transit.jl:702-765) of nvec host vectors (nvec x mdim) with J values
 * totJs[n] (the state's J, not 2J; J < 0 is skipped): mu[n] in nuclear magnetons and
 * Q[n] in e fm^2, NaN where the reference prints nothing (J < 0, or a vanishing 3j
 * symbol at the basis' M, e.g. mu of any state of an M = 0 basis). */
int32_t smgpu_eval_moment(smgpu_moments m, int32_t nvec, const double* vecs, const double* totJs, double* mu,
                          double* Q);
/* the oscillator length in use (fm) */
int32_t smgpu_moments_bpar(smgpu_moments m, double* bpar);

/* -------------------------------------------------------------- transitions
 * transit_main (transit.jl:561-700): M1 and E2 transitions between the states of
 * two problems on one interaction and nucleus, the left at 2M_l and the right at
 * 2M_r <= 2M_l (transit_main puts each at its 2J).  smgpu_transitions_create
 * builds prep_obtr_nondiag's one-body operators (transit.jl:286-445), the
 * components mu = (2M_l - 2M_r) / 2 <= 2 of M1 and E2 from the right basis into
 * the left one, and uploads them to `device`; gfactors, effcharge and bpar as
 * smgpu_moments_create. */
int32_t smgpu_transitions_create(smgpu_problem left, smgpu_problem right, int32_t device, const double* gfactors,
                                 const double* effcharge, double bpar, smgpu_transitions* out);
int32_t smgpu_transitions_destroy(smgpu_transitions t);
/* w_left = O v_right with O the M1 (which = 0) or E2 (1) operator (op_obtr_nondiag,
 * transit.jl:447-499), FP64 device vectors, on `stream` (NULL: the handle's own) */
int32_t smgpu_transitions_apply(smgpu_transitions t, int32_t which, const double* v_right_dev, double* w_left_dev,
                                void* stream);
/* the same on host vectors: nvec right vectors (mdim_r each) to nvec left ones (mdim_l each) */
int32_t smgpu_transitions_apply_host(smgpu_transitions t, int32_t which, int32_t nvec, const double* v_right,
                                     double* w_left);
/* transit_main's reduced matrix elements Mred(l, r) = <l|O|r> / 3j(J_l lam J_r; -M_l mu M_r)
 * of nl left (nl x mdim_l) and nr right (nr x mdim_r) host vectors with 2J = jl2 and jr2:
 * m1[l * nr + r] (lam 1, mu_N) and e2[l * nr + r] (lam 2, e fm^2), NaN where the
 * transition is not allowed or the 3j symbol vanishes.  B(lam; r -> l) = Mred^2 / (jr2 + 1). */
int32_t smgpu_transit(smgpu_transitions t, int32_t jl2, int32_t jr2, int32_t nl, const double* wl, int32_t nr,
                      const double* wr, double* m1, double* e2);
/* 2M_l - 2M_r, the oscillator length in use (fm), and the two basis dimensions */
int32_t smgpu_transitions_info(smgpu_transitions t, int32_t* mu2, double* bpar, int64_t* mdim_l, int64_t* mdim_r);

/* ------------------------------------------------------------------ serving
 * Persistent request loop (no counterpart in the reference, whose main_sm
 * solves one nucleus per call; each request is one main_sm eigenvalue solve,
 * shellmodel_main.jl:89-281, with the observable en[1]).  Requests are prepared
 * on host threads (readsmsnt through the table build), admitted in arrival
 * order as device slots and HBM free up (a smaller later request may pass a
 * waiting larger one a bounded number of times), and run concurrently, one
 * HIP stream per slot.  Requests with equal orbits, valence numbers and 2M share
 * one device copy of the index tables; each keeps its own coefficients and
 * Lanczos basis.  All calls are thread-safe. */
#define SMGPU_ETIMEDOUT (-7)

#define SMGPU_REQ_QUEUED 0     /* waiting for a host thread */
#define SMGPU_REQ_PREPARING 1  /* host tables being built */
#define SMGPU_REQ_READY 2      /* waiting for a slot / HBM */
#define SMGPU_REQ_RUNNING 3
#define SMGPU_REQ_DONE 4
#define SMGPU_REQ_FAILED 5
#define SMGPU_REQ_CANCELED 6

typedef struct {
    int32_t device;
    int32_t max_inflight;   /* concurrent solves (streams); <= 0: 4 */
    int64_t hbm_budget;     /* HBM bytes the server may hold; <= 0: 90% of free memory at create */
    int32_t host_threads;   /* requests prepared concurrently; <= 0: 2 */
    int32_t table_threads;  /* threads per table build; <= 0: cores / host_threads */
    int64_t keep_bytes;     /* an idle slot frees its buffers above this; <= 0: 16 GiB */
    int32_t max_batch;      /* requests on one basis and precision that one slot solves in
                               lockstep with shared H·v launches; <= 0: 1 (no batching) */
    int32_t batch_variants; /* nonzero: a batch also takes requests with other interactions on
                               its basis (each member with its own coefficients, H·v launches
                               sharing the index loads); 0: identical Hamiltonians only.
                               smgpu_default_server_config sets 1 */
    int64_t large_mdim;     /* > 0: requests of at least this basis dimension run one solve
                               (or batch) at a time, oldest first; the others run smallest
                               basis first (oldest first once they have waited 60 s), and no
                               request holds up another.  0: arrival order, a request that
                               does not fit yet bypassed by at most 4 * max_inflight others */
    int32_t reserve_cus;    /* with large_mdim and max_inflight >= 2: slot 0 runs the large
                               requests only; while smaller requests are in the server, on
                               streams that leave this many CUs (at most half the device) to
                               the other slots, otherwise on the whole device; 0: no CU
                               partition */
    int32_t reserved_;      /* zero */
} smgpu_server_config;

typedef struct {
    int32_t state;          /* SMGPU_REQ_* */
    int32_t status;         /* SMGPU_OK or the failure code once finished */
    int32_t num_ev;
    int32_t index_shared;   /* 1: device index tables were already resident */
    int32_t slot;           /* slot that ran it (-1 before admission) */
    int32_t lm;             /* Krylov basis size it runs with (the request's lm, or the
                               largest that fits the HBM budget; set once prepared) */
    int64_t mdim;
    int64_t footprint;      /* HBM bytes the request needs (index + coefficients + workspace) */
    int64_t index_bytes;    /* of which the shared index tables */
    double submit_ms;       /* times on the server clock (ms since create) */
    double prep_start_ms, prep_end_ms, admit_ms, done_ms;
    double latency_ms;      /* done - submit */
    smgpu_solve_stats solve;
} smgpu_request_info;

typedef struct {
    int64_t submitted, completed, failed, canceled;
    int32_t queued, preparing, ready, running;
    int32_t peak_running, slots;
    int64_t hbm_budget, hbm_in_use, hbm_peak;
    int64_t index_hits, index_misses, index_uploads, index_evictions;
    double clock_ms;        /* server clock now */
    int64_t batched;        /* requests that ran in a batch of two or more */
    int32_t peak_batch;     /* largest batch run */
    int32_t reserved_;
} smgpu_server_stats;

void smgpu_default_server_config(smgpu_server_config* cfg);
int32_t smgpu_server_create(const smgpu_server_config* cfg, smgpu_server* out);
/* Cancels requests not yet finished, waits for the slots, frees everything. */
int32_t smgpu_server_destroy(smgpu_server s);
/* One main_sm solve: interaction file, nucleus ("Cr48"), 2M (< 0: A % 2) and
 * solver parameters (num_ev <= 64).  Returns at once with the request id.  A
 * request whose Lanczos basis at params->lm does not fit the server's HBM budget
 * runs with the largest lm that does (not below max(ls + 20, 2 ls)); info.lm
 * reports it. */
int32_t smgpu_server_submit(smgpu_server s, const char* snt_path, const char* nucleus, int32_t twoM,
                            const smgpu_solve_params* params, int64_t* request_id);
int32_t smgpu_server_poll(smgpu_server s, int64_t request_id, int32_t* state);
/* Block until the request has finished (done, failed or canceled) or timeout_ms
 * (< 0: no limit) passes: SMGPU_OK or SMGPU_ETIMEDOUT. */
int32_t smgpu_server_wait(smgpu_server s, int64_t request_id, double timeout_ms);
/* Finished request: evals[num_ev] (may be NULL) and info (may be NULL); the
 * record is dropped afterwards.  Returns the request's own status. */
int32_t smgpu_server_result(smgpu_server s, int64_t request_id, double* evals, smgpu_request_info* info);
/* Info of a request in any state (the record stays). */
int32_t smgpu_server_info(smgpu_server s, int64_t request_id, smgpu_request_info* info);
int32_t smgpu_server_cancel(smgpu_server s, int64_t request_id);
int32_t smgpu_server_get_stats(smgpu_server s, smgpu_server_stats* st);

/* Device memory helpers for callers that keep vectors resident. */
int32_t smgpu_malloc(smgpu_ham h, int64_t bytes, void** dev);
int32_t smgpu_free(smgpu_ham h, void* dev);
int32_t smgpu_memcpy(smgpu_ham h, void* dst, const void* src, int64_t bytes, int32_t kind /*0 H2D,1 D2H,2 D2D*/);
int32_t smgpu_synchronize(smgpu_ham h);

/* Self-test of the shared tridiagonal eigen-solver (host build of the same
 * source the device kernels use): lowest k eigenpairs of T. */
int32_t smgpu_selftest_tridiag(const double* alpha, const double* beta, int32_t n, int32_t k, double* vals,
                               double* vecs);

/* Self-test of the host dense symmetric eigen-solver of the block Lanczos matrix
 * (Householder reduction + tridiagonal bisection / inverse iteration): lowest k
 * eigenvalues of the n x n row-major symmetric A, and their unit eigenvectors
 * (vector j at vecs + j*n) unless vecs is NULL.  Without vecs, a banded A (lower
 * bandwidth below n / 4) goes through the Givens band-to-tridiagonal reduction the
 * solver uses for its per-step eigenvalues. */
int32_t smgpu_selftest_symeig(const double* A, int32_t n, int32_t k, double* vals, double* vecs);

/* Self-test of the thick-restart small problem (tri_restart) for the tridiagonal
 * T(alpha[m], beta[m]; beta[m-1] = coupling to the residual vector): the rotation
 * S[ls*m] (column j at S + j*m) and the new leading block alpha_out[ls],
 * beta_out[ls].  device < 0 runs the host build, otherwise the device kernel. */
int32_t smgpu_selftest_restart(int32_t device, const double* alpha, const double* beta, int32_t m, int32_t ls,
                               double* S, double* alpha_out, double* beta_out);

#ifdef __cplusplus
}
#endif
#endif
