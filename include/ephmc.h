/*
 * ephmc.h -- C-ABI of the MI355X electron-phonon HMC engine (libephmc.so).
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.  Every pointer
 * argument named *_dev is a device (HBM) pointer; every *_host pointer is host memory.
 * Matrices are N x N column-major, i.e. the reference's row-major chunks read as LAPACK
 * columns (organic_dense_matrix.c:580).  Fields x, p, p_dot use the reference layout
 * x[tau*3N + site*3 + direction] (organic_hmc.c:195-197).
 *
 * One context = one chain on one GPU, bound to one HIP stream.  Calls on a context are
 * asynchronous on its stream unless they return a host value (documented per call).
 * Contexts are independent; one context must not be used from two threads at once.
 *
 * Each entry point names the reference function it replaces (file:line in
 * j-ostmeyer/electron-phonon-hmc, directory simulation/).
 */
#ifndef EPHMC_H
#define EPHMC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPHMC_OK              0
#define EPHMC_ERR_ARG        -1   /* invalid argument or size                      */
#define EPHMC_ERR_HIP        -2   /* HIP runtime error (message in ephmc_last_error) */
#define EPHMC_ERR_UNSUPPORTED -3  /* flag combination not provided by this build    */
#define EPHMC_ERR_ALLOC      -4   /* device allocation failed                       */
#define EPHMC_ERR_NUMERIC    -5   /* singular matrix (zero pivot)                   */

#define EPHMC_FP64  0   /* plain FP64 drop-in path                                     */
#define EPHMC_MIXED 1   /* FP32 slice products and force in the molecular dynamics;
                           FP64 factorisations, log-det and dH of the Metropolis step */

typedef struct ephmc_ctx ephmc_ctx;

/* Model and run configuration.  Couplings are in lattice units, i.e. after the scaling of
 * organic_main.c:99-116 (t *= beta/Nt, kappa = lambda*t/x_avg, mu = mu0*beta/Nt - 2 sum|t|,
 * omega *= beta/Nt).  Flags mirror organic_flags (organic_flags.h:4-14). */
typedef struct ephmc_config {
	unsigned l1, l2, nt;
	double t[3], kappa[3], mu, omega;
	unsigned flavors;
	unsigned steps;          /* molecular-dynamics steps per trajectory (Nmd)            */
	double traj_length;      /* 0 selects pi/2 (organic_main.c:114-116)                  */
	int taylor_order;        /* 2, 4 (reference default), 6 or 8                          */
	int precision;           /* EPHMC_FP64 or EPHMC_MIXED                                */
	unsigned checkpoint;     /* slices between stored partial products in the force (k);
	                            k = 1 stores every slice and the force sweep reads them back
	                            instead of recomputing; 0: k = 1 while nt N^2 values fit in
	                            4 GiB, else 16                                               */
	int zero_filling, single_particle, constant_force, no_fourier_acc, classical_fourier_acc;
	unsigned cfa_steps;
	double cfa_mass;
	int device;              /* HIP device ordinal                                       */
	void *stream;            /* hipStream_t to run on; NULL creates a default-priority one */
	int graph;               /* 1: ephmc_trajectory replays its molecular dynamics as a
	                            captured hipGraph (captured on first use; not while timing) */
	int schur_form_inv;      /* SCHUR_FORM_INV: every M_hat^{-1} and log det through the real Schur
	                            form of M_hat - 1 with small entries zeroed
	                            (organic_dense_matrix.c:556-603), bond projections clipped to
	                            [-2, 2] (organic_hmc.c:113-114), CONSTANT_FORCE with
	                            QM1P = 1 - M_hat^{-1} (organic_dense_matrix.c:270-278) */
	unsigned stabilize;      /* k > 0: M_hat^{-1} and its log det (Metropolis and molecular
	                            dynamics) from the slice product kept as U diag(d) V with an FP64
	                            QR every k slices (ephmc_stable_logdet_inv); 0: the plain product */
} ephmc_config;

/* Per-trajectory record (organic_hmc.c:233-245, results columns 3,4,7,8). */
typedef struct ephmc_traj_record {
	int accepted;
	double boltzmann;        /* exp(H_old - H_new)                      */
	double energy;           /* H after the Metropolis decision         */
	double logdet;           /* log det(1 + B..B) after the decision    */
	double dH;               /* H_new - H_old of the proposal           */
} ephmc_traj_record;

/* ---- lifecycle ---------------------------------------------------------------------- */
/* EPHMC_ERR_ARG for an empty lattice, omega zero or NaN, or an unknown precision, before any device
 * call; EPHMC_ERR_UNSUPPORTED beyond 2048 sites or for excluded flag combinations. */
int ephmc_create(const ephmc_config *cfg, ephmc_ctx **out);
int ephmc_destroy(ephmc_ctx *ctx);
int ephmc_synchronize(ephmc_ctx *ctx);
const char *ephmc_last_error(void);
int ephmc_dims(const ephmc_ctx *ctx, unsigned *ns, unsigned *nt, size_t *dim);
/* neighbour table nnt[12N] of construct_triangular (organic_lattice.c:10-46) */
int ephmc_lattice(const ephmc_ctx *ctx, unsigned *nnt_host);

/* ---- device memory helpers (for hosts that do not include HIP headers) --------------- */
int ephmc_malloc(size_t bytes, void **dev);
int ephmc_free(void *dev);
int ephmc_memcpy_h2d(ephmc_ctx *ctx, void *dev, const void *host, size_t bytes);  /* synchronous */
int ephmc_memcpy_d2h(ephmc_ctx *ctx, void *host, const void *dev, size_t bytes);  /* synchronous */

/* ---- (a) slice exponential ----------------------------------------------------------- */
/* M <- p(A_tau) M with the pair factors [0, stop) (revert = 0) or np-1 down to stop+1
 * (revert != 0); replaces apply_exp_mu_K_to_M (organic_dense_matrix.c:209-216).
 * x_tau_dev: 3N couplings of one slice; M_dev: N x N, FP64. */
int ephmc_apply_exp_M(ephmc_ctx *ctx, const double *x_tau_dev, double *M_dev, int revert, int stop);

/* ---- (b) chained slice product -------------------------------------------------------- */
/* M_hat = 1 + B_0 B_1 ... B_{Nt-1} (no +1 for ZERO_FILLING / SINGLE_PARTICLE);
 * replaces construct_M_hat (organic_dense_matrix.c:251-265).  FP64. */
int ephmc_construct_M_hat(ephmc_ctx *ctx, const double *x_dev, double *M_hat_dev);

/* ---- (c) log-det, inverse, force ------------------------------------------------------ */
/* Stabilised log|det(1 + B_0 ... B_{nt-1})| and its inverse for field x: the product grown k slices
 * at a time in the slice precision (EPHMC_FP64, or EPHMC_MIXED for FP32 slices in a mixed context),
 * kept as U diag(d) V through a pre-pivoted FP64 Householder QR after every block, and inverted as
 * X^{-1} D_b^{-1} U^T with X = D_b^{-1} U^T + D_s V (D_b = max(d, 1), D_s = min(d, 1)).  The
 * periodic stabilisation that construct_M_hat + logDetM_and_inv_M_hat (organic_dense_matrix.c:251-265,
 * :547-607) do without: equal to them where the plain product is well conditioned, and accurate
 * where it is not.  M_inv_dev: N x N.  logdet_host / sign_host may be NULL (then asynchronous). */
int ephmc_stable_logdet_inv(ephmc_ctx *ctx, const double *x_dev, int precision, unsigned k, double *M_inv_dev,
		double *logdet_host, int *sign_host);
/* log|det M_hat| and M_hat <- M_hat^{-1} in place (FP64, partial pivoting); replaces
 * logDetM_and_inv_M_hat (organic_dense_matrix.c:547-607).  SINGLE_PARTICLE: log tr M_hat,
 * M_hat unchanged.  logdet_host / sign_host may be NULL; when given, the call waits and
 * reports a singular matrix (a zero pivot) as EPHMC_ERR_NUMERIC with *logdet_host = -inf
 * (the reference's log|0|, :582-584) and M_hat filled with NaN. */
int ephmc_logdet_inv(ephmc_ctx *ctx, double *M_hat_dev, double *logdet_host, int *sign_host);
/* Real Schur form B = Z T Z^T of an N x N matrix (column-major, device): T quasi-upper-triangular
 * with 2 x 2 blocks in standard form (real pair: triangular; complex pair: equal diagonal, off-
 * diagonal entries of opposite sign), Z orthogonal, eigenvalues wr + i wi (host arrays, may be
 * NULL); the factorisation of LAPACK dgees('V', 'N') that the SCHUR_FORM_INV branch uses
 * (organic_dense_matrix.c:560).  Hessenberg reduction and Francis double-shift QR, FP64.  Waits;
 * EPHMC_ERR_NUMERIC if the iteration does not converge. */
int ephmc_schur(ephmc_ctx *ctx, const double *B_dev, double *Z_dev, double *T_dev, double *wr_host, double *wi_host);
/* Row interchanges of the last ephmc_logdet_inv in LAPACK's 1-based ipiv form (row i was
 * exchanged with row ipiv[i] at step i): the identity when no exchange was needed, else the
 * pivoted factorisation's; the p argument of logDetM_and_inv_M_hat (dgetrf's ipiv,
 * organic_dense_matrix.c:580).  ipiv_host: N ints.  Waits. */
int ephmc_last_pivots(ephmc_ctx *ctx, int *ipiv_host);
/* p_dot = fermion force for field x given M_inv = (1 + B..B)^{-1}; replaces force
 * (organic_hmc.c:86-134) with construct_all_QM1P + refine_QM1P
 * (organic_dense_matrix.c:311-442) computed as an adjoint sweep.  Uses the context's
 * precision and flags: CONSTANT_FORCE puts the slice-0 force (construct_QM1P at time 0,
 * :267-309) on every slice; SINGLE_PARTICLE ignores M_inv (may be NULL) and scales by
 * exp(-logdet); NO_FOURIER_ACC adds the boson force (organic_bosons.c:179-194). */
int ephmc_force(ephmc_ctx *ctx, const double *x_dev, const double *M_inv_dev, double logdet, double *p_dot_dev);

/* ---- retained dense-matrix entry points (organic_dense_matrix.h:14-38) --------------------
 * FP64, on the context's lattice; x_tau_dev: the 3N couplings of one slice; a "column" is one
 * contiguous chunk of N values (the reference's LAPACK column). */
/* out = A_tau in for ncols columns, A_tau = mu + K(x_tau) (out of place); replaces
 * apply_mu_K_to_v (organic_dense_matrix.c:96-120; ncols = 1) and apply_mu_K_to_M (:123-131;
 * ncols = N, with the result copied back) */
int ephmc_apply_mu_K(ephmc_ctx *ctx, const double *x_tau_dev, const double *in_dev, double *out_dev, unsigned ncols);
/* M <- prod_j (1 + a_lin[j] A_tau + a_quad[j] A_tau^2) M for ncols <= N columns, j = 0..npairs-1
 * in order (a_quad[j] = 0: the linear factor); a_lin / a_quad are host arrays.  Replaces
 * apply_2nd_order_mu_K_to_v (:168-183; one pair, a_lin = 2 Re a, a_quad = |a|^2 or 0 for real a)
 * and apply_K_plus_1_v (:160-165; a_lin = 1) */
int ephmc_apply_pairs(ephmc_ctx *ctx, const double *x_tau_dev, double *M_dev, unsigned ncols, int npairs,
		const double *a_lin_host, const double *a_quad_host);
/* M <- partial pair product of the context's Taylor pairs for ncols <= N columns: positions
 * [0, stop) (revert = 0) or np-1 down to stop+1 (revert != 0); replaces
 * apply_pair_expansion_mu_K_to_v / apply_exp_mu_K_to_v (:186-207) */
int ephmc_apply_exp_cols(ephmc_ctx *ctx, const double *x_tau_dev, double *M_dev, unsigned ncols, int revert, int stop);
/* Q <- sum_j a_lin_j Z_j + a_quad_j (A Z_j + Z_j A), Z_j = P_<j Q P_>j: the derivative of the
 * pair product of one slice (tr[Q dB] = tr[Y dA]); replaces refine_QM1P (:383-442) */
int ephmc_refine_QM1P(ephmc_ctx *ctx, const double *x_tau_dev, double *Q_dev);
/* Q = refine(B_{time+1} ... B_{nt-1} M_inv B_0 ... B_{time-1}) (identity for M_inv under
 * SINGLE_PARTICLE, M_inv may then be NULL); replaces construct_QM1P (:267-309, the non-Schur
 * branch) */
int ephmc_construct_QM1P(ephmc_ctx *ctx, const double *x_dev, const double *M_inv_dev, double *Q_dev, int time);
/* Q[tau] = B_{tau+1} ... B_{nt-1} M_inv B_0 ... B_{tau-1} for every tau (Q: nt N^2 doubles), by
 * recursive halving in nt log2(nt) slice applications; replaces construct_all_QM1P (:311-381) */
int ephmc_construct_all_QM1P(ephmc_ctx *ctx, const double *x_dev, const double *M_inv_dev, double *Q_dev);

/* One molecular-dynamics force at field x exactly as the leapfrog computes it: the forward chain
 * in the context's precision (FP32 under EPHMC_MIXED) gives M_hat, the FP64 Gauss-Jordan its
 * inverse and log det, the adjoint sweep the force; replaces the construct_M_hat +
 * logDetM_and_inv_M_hat + force sequence of leap_frog (organic_hmc.c:178-183). */
int ephmc_md_force(ephmc_ctx *ctx, const double *x_dev, double *p_dot_dev);

/* ---- (d) Fourier-accelerated bosons ---------------------------------------------------- */
/* x = free-boson sample coloured from dim normals; replaces sample_free_bosons
 * (organic_bosons.c:55-74) after its random_vector draw. */
int ephmc_sample_free_bosons(ephmc_ctx *ctx, const double *normals_dev, double *x_dev);
/* p = Fourier-accelerated momenta coloured from dim normals; replaces
 * sample_fourier_momenta (organic_bosons.c:76-100) after its random_vector draw. */
int ephmc_sample_fourier_momenta(ephmc_ctx *ctx, const double *normals_dev, double *p_dev);
/* kinetic energy of p; replaces energy_momenta / energy_fourier_momenta
 * (organic_hmc.c:56-65, organic_bosons.c:102-124).  Waits for the result. */
int ephmc_energy_momenta(ephmc_ctx *ctx, const double *p_dev, double *energy_host);
/* exact harmonic flow of every Matsubara mode for time h; replaces evolve_bosons
 * (organic_bosons.c:126-150) */
int ephmc_evolve_bosons(ephmc_ctx *ctx, double *x_dev, double *p_dev, double h);
/* drift of the position update with the context's momentum scheme: Fourier-accelerated flow
 * for time |half| h, x += |half| h p (NO_FOURIER_ACC) or the CLASSICAL_FOURIER_ACC sub-steps
 * with the boson force; replaces update_x (organic_hmc.c:136-165) */
int ephmc_update_x(ephmc_ctx *ctx, double *x_dev, double *p_dev, double h, double half);
/* p_dot (+)= -dS_B/dx = -omega x + (x_{tau+1} - 2 x_tau + x_{tau-1}) / omega, overwriting
 * p_dot when overwrite != 0; replaces boson_force (organic_bosons.c:179-194) */
int ephmc_boson_force(ephmc_ctx *ctx, const double *x_dev, double *p_dot_dev, int overwrite);
/* x += h M^{-1} p with the Fourier-space mass of the context's kinetic term (omega ->
 * sqrt(cfa_mass^2 + omega^2) under CLASSICAL_FOURIER_ACC); replaces evolve_bosons_approx_step
 * (organic_bosons.c:152-177) */
int ephmc_evolve_bosons_approx_step(ephmc_ctx *ctx, double *x_dev, const double *p_dev, double h);
/* H = -flavors logdet + energyP + S_B(x); replaces hamilton (organic_hmc.c:67-84).  Waits. */
int ephmc_hamilton(ephmc_ctx *ctx, const double *x_dev, double logdet, double energyP, double *H_host);
/* position-Verlet leapfrog of Nmd steps on x, p (the context's precision); replaces
 * leap_frog (organic_hmc.c:167-190) */
int ephmc_leap_frog(ephmc_ctx *ctx, double *x_dev, double *p_dev);

/* ---- HMC chain with device-resident state ---------------------------------------------- */
/* Initial field and M_hat^{-1}: replaces run_hmc's start (organic_hmc.c:325-335).
 * normals_host: dim normals of the free-boson draw. */
int ephmc_chain_init(ephmc_ctx *ctx, const double *normals_host);
/* One trajectory (organic_hmc.c:192-243): momenta from dim normals, leapfrog, FP64
 * M_hat / log-det / H, Metropolis with the given uniform (the reference's genrand64_real2
 * draw that follows the normals).  Asynchronous when rec == NULL; with rec the call waits. */
int ephmc_trajectory(ephmc_ctx *ctx, const double *normals_host, double uniform, ephmc_traj_record *rec);
/* Record of the last trajectory (waits). */
int ephmc_last_record(ephmc_ctx *ctx, ephmc_traj_record *rec);
/* Copies of the chain state (wait): x (dim), M_hat^{-1} (N*N). */
int ephmc_chain_get_x(ephmc_ctx *ctx, double *x_host);
int ephmc_chain_set_x(ephmc_ctx *ctx, const double *x_host);   /* also rebuilds M_hat^{-1} */
int ephmc_chain_get_minv(ephmc_ctx *ctx, double *minv_host);
int ephmc_chain_get_logdet(ephmc_ctx *ctx, double *logdet_host);
/* device pointers of the chain state (valid until ephmc_destroy) */
int ephmc_chain_buffers(ephmc_ctx *ctx, double **x_dev, double **p_dev, double **minv_dev);

/* ---- many chains in one launch per phase ------------------------------------------------- *
 * The reference runs one chain per process (organic_hmc.c:296-353, run_hmc).  A batch advances n
 * chains (contexts after ephmc_chain_init) by one trajectory each, with each kernel of the
 * trajectory launched once over all chains of a group: chains with the same arithmetic, force
 * columns per workgroup, MD steps and workgroup class (N <= 256 / 512 / 1024).  Runs the
 * Fourier-accelerated trajectory (no NO_FOURIER_ACC / CLASSICAL_FOURIER_ACC, stabilize,
 * SINGLE_PARTICLE, ZERO_FILLING, CONSTANT_FORCE or SCHUR_FORM_INV) at Taylor order 4 for
 * 33 <= N <= 1024, else EPHMC_ERR_UNSUPPORTED.  Each chain's records and state are bit for bit
 * those of ephmc_trajectory with the same inputs.  ephmc_batch_create sets each context's force
 * sweep to four columns per workgroup where its LDS allows (the width that runs many chains
 * fastest; ephmc_autotune's choice is for a chain alone), and the contexts keep it.  The contexts
 * keep their state: between batch trajectories (after ephmc_batch_sync) each can be used alone.  While a batch holds a context,
 * ephmc_destroy and ephmc_autotune refuse it (EPHMC_ERR_ARG): destroy the batch first. */
typedef struct ephmc_batch ephmc_batch;
int ephmc_batch_create(ephmc_ctx *const *ctxs, unsigned n, ephmc_batch **out);
/* One trajectory of every chain: normals[i] (dim_i values) and uniforms[i] as for
 * ephmc_trajectory.  Asynchronous (the inputs are copied before it returns). */
int ephmc_batch_trajectory(ephmc_batch *b, const double *const *normals, const double *uniforms);
/* Waits for the trajectory in flight; recs (n records, may be null) receive its records. */
int ephmc_batch_sync(ephmc_batch *b, ephmc_traj_record *recs);
/* Number of groups and of kernel launches per batch trajectory. */
int ephmc_batch_info(const ephmc_batch *b, unsigned *groups, unsigned *launches);
int ephmc_batch_destroy(ephmc_batch *b);

/* ---- measurement -------------------------------------------------------------------------- */
/* Current-current correlator g_ab(tau) (nt x 3 x 3, tau-major) of field x with
 * M_inv = (1 + B..B)^{-1}; replaces cc_correlator (organic_dense_matrix.c:451-524).  FP64.
 * Allocates 2 nt N^2 doubles of device memory on first use.  Waits. */
int ephmc_cc_correlator(ephmc_ctx *ctx, const double *x_dev, const double *M_inv_dev, double logdet, double *greens_host);

/* Allocation-free variant of ephmc_cc_correlator.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops).  The
 * transfer matrices and block sums live in a static 64 MiB device array (one per GPU, held until
 * the process ends), so 2 nt N^2 + 9 nt ceil(N/8) <= 8,388,608 doubles (e.g. N = 256, Nt <= 63;
 * larger sizes return EPHMC_ERR_ARG).  greens_dev: nt x 9 doubles in device memory, written
 * asynchronously on the context's stream (no host copy, no wait).  Every call on one GPU shares
 * that array: calls that overlap on one GPU (two streams or threads) race. */
int ephmc_cc_correlator_static(ephmc_ctx *ctx, const double *x_dev, const double *M_inv_dev, double logdet,
		double *greens_dev);

/* ---- (e) transient localisation ---------------------------------------------------------- */
/* Batched disorder samples; replaces hist_trans_loc x meas (organic_trans_loc.c:25-83,
 * :101-105) with the normals of each sample given (3N per sample, sample-major).
 * greens_host[nbins] is accumulated (+=, not yet divided by the sample count);
 * results_host[nsamples] receives sum exp(-E)/N per sample.  contr_mat: 2 x 3. */
int ephmc_tl_hist(ephmc_ctx *ctx, const double *normals_host, unsigned nsamples,
		const double *contr_mat_host, double width, unsigned nbins,
		double *greens_host, double *results_host);

/* ---- launch-shape tuning ----------------------------------------------------------------------
 * No counterpart in the reference (its shapes are fixed by the CPU loops).  Times the forward
 * chain (columns per workgroup x sites per thread of a lattice row) and the force sweep (columns
 * per workgroup) of this context's size and precision for every shape that fits LDS, and in
 * mixed precision also the fp64 chain of the Metropolis log det; keeps the fastest of each for
 * the context's later calls, and
 * drops a captured trajectory graph so that it is captured again.  Waits; call it between
 * trajectories.  res (may be null) receives the choice and its per-call times. */
typedef struct ephmc_tune_result {
	unsigned chain_cols, adjoint_cols;
	double chain_ms, adjoint_ms;
	int adjoint_w0_early;    /* 1: the sweep writes each slice's W_0 ahead of its first phase */
	unsigned chain_strip;    /* sites per thread of a lattice row in the chain (0: one site) */
	unsigned chain64_cols, chain64_strip;   /* fp64 chain of the Metropolis log det (mixed precision; */
	double chain64_ms;                      /* in fp64 contexts the same as chain_*)               */
} ephmc_tune_result;
int ephmc_autotune(ephmc_ctx *ctx, ephmc_tune_result *res);

/* ---- diagnostics ---------------------------------------------------------------------------- */
/* Number of inversions that needed row exchanges (a multiplier above 1 in the elimination
 * without exchanges) and ran the pivoted fallback; waits. */
int ephmc_gj_fallbacks(ephmc_ctx *ctx, unsigned long *count);
/* Average duration (ms, HIP events on the context's stream) of the named kernel group since
 * the last reset; groups: "chain", "gj", "adjoint", "reduce", "fft", "tl".  Timing is only
 * recorded after ephmc_timing(ctx, 1). */
int ephmc_timing(ephmc_ctx *ctx, int enable);
int ephmc_timing_get(ephmc_ctx *ctx, const char *group, double *avg_ms, unsigned long *calls);

#ifdef __cplusplus
}
#endif
#endif
