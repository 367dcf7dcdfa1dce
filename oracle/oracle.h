/*
 * oracle.h -- CPU restatement of the electron-phonon HMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the C host
 * layer, organic_main) links, loads or calls this code.  It is used by tests/,
 * by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg
 * (timed as the "port" CPU baseline).
 *
 * Every function restates the algorithm of the reference (j-ostmeyer/electron-phonon-hmc,
 * directory simulation/) and cites the file:line it follows.  The arithmetic order
 * follows the reference wherever that is practical (scatter-order stencil, pair
 * expansion, binary-splitting force), so the oracle is also a faithful CPU timing
 * baseline of the reference algorithm.
 *
 * How each function is pinned (tests/test_oracle_ref.py, test_oracle_math.py, test_oracle_pins.py):
 *   orc_mt_*, orc_random_vector, orc_construct_triangular, orc_construct_contraction_mat,
 *     orc_contract_greens, orc_fprint_results, orc_fprint_corr: bit-exact against the reference's
 *     own sources compiled into oracle/_ref/libref.so, plus the C++ standard's MT19937-64 answer;
 *   orc_set_taylor: every table's pair product equals sum_{k<=order} z^k/k! (orders 2, 4, 6, 8),
 *     and partial products follow the hand-derived application order;
 *   orc_scale_params: closed-form numbers of organic_main.c:99-116 (both branches);
 *   orc_apply_mu_K_to_v/M, orc_apply_exp_mu_K_to_v/M, orc_construct_M_hat, orc_transpose: dense
 *     numpy brute force, and the closed form for spatially uniform fields;
 *   orc_logdet_inv: numpy slogdet / inv;
 *   orc_construct_all_QM1P: dense numpy products; orc_refine_QM1P, orc_force,
 *     orc_construct_QM1P: central finite differences of the log det, the uniform-field closed
 *     form, and the constant-force slice-0 identity;
 *   orc_r2c_many / orc_c2r_many: numpy rfft / irfft (FFTW's unnormalised half spectrum);
 *   orc_sample_*, orc_energy_momenta, orc_evolve_bosons: numpy colouring, the full-spectrum
 *     energy and the explicit mode rotation (energy-conserving, h -> -h inverse);
 *   orc_evolve_bosons_approx_step: the gradient of orc_energy_momenta with the CFA mass and a
 *     numpy rfft restatement; orc_boson_force: central differences of the boson action;
 *   orc_update_x, orc_leap_frog: second-order convergence of dH and exact p -> -p reversibility
 *     in the Fourier-accelerated, NO_FOURIER_ACC and CLASSICAL_FOURIER_ACC schemes;
 *   orc_hamilton, orc_higher_matsubaras_norm: numpy sums; orc_trajectory: kappa = 0 exactness,
 *     <exp(-dH)> = 1 and its results columns recomputed from the state;
 *   orc_cc_correlator: dense numpy evaluation from explicit transfer matrices and the
 *     plane-wave closed form for fields uniform in space;
 *   orc_construct_hamilton, orc_apply_current_to_v, orc_eigh, orc_hist_trans_loc, orc_tl_width:
 *     numpy eigh, brute-force binning and the closed-form bin width.
 */
#ifndef EPHMC_ORACLE_H
#define EPHMC_ORACLE_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: MT19937-64 (simulation/mt19937-64.c:73-161) + Box-Muller (organic_random.c:11-22) */
typedef struct { uint64_t s[312]; int i; } orc_mt;
void orc_mt_seed(orc_mt *m, uint64_t seed);
uint64_t orc_mt_u64(orc_mt *m);
double orc_mt_real2(orc_mt *m);
void orc_random_vector(orc_mt *m, double *x, unsigned n);

/* ---- lattice (simulation/organic_lattice.c:10-106) */
unsigned *orc_construct_triangular(unsigned l1, unsigned l2, unsigned *ns, unsigned *nn);
double *orc_construct_contraction_mat(const double *a, const double *b, unsigned n);
double orc_contract_greens(const double *gr, const double *mat, unsigned nn, unsigned tau);

/* ---- flags (simulation/organic_flags.h:4-14) */
typedef struct {
	int schur_form_inv, zero_filling, static_disorder, single_particle;
	int constant_force, no_fourier_acc, classical_fourier_acc;
	unsigned cfa_steps;
	double cfa_mass;
} orc_flags;

/* ---- model: lattice-unit parameters after organic_main.c:99-116 */
#define ORC_MAX_PAIRS 26
typedef struct {
	unsigned ns, nn, nt;
	const unsigned *nnt;
	double t[3], kappa[3], mu, omega;
	unsigned flavors;
	int np;                          /* number of Taylor pairs                          */
	double a_re[ORC_MAX_PAIRS];      /* pair coefficient a_i = .5*c[order[i]]/np, i = application position */
	double a_im[ORC_MAX_PAIRS];
	orc_flags mode;
} orc_model;

/* Taylor pair tables (simulation/taylor-coeff_pairs.h) and ordering
 * (set_taylor_pair_order, organic_dense_matrix.c:133-159).  which: 4 = TAYLOR_ORDER_4 (default). */
int orc_set_taylor(orc_model *m, int which);

/* parameter scaling organic_main.c:99-116 */
void orc_scale_params(double beta_in, const double *J, const double *lambda, double mu0, double omega0,
		unsigned nt, int static_disorder, double *t, double *kappa, double *mu, double *omega, double *x_avg_out);

/* ---- slice stencil and exponential (organic_dense_matrix.c:96-216) */
void orc_apply_mu_K_to_v(const orc_model *m, const double *x, const double *v, double *w, double mu);
void orc_apply_mu_K_to_M(const orc_model *m, const double *x, double *M, double *w);
void orc_apply_exp_mu_K_to_v(const orc_model *m, const double *x, double *v, double *w, int revert, int stop);
void orc_apply_exp_mu_K_to_M(const orc_model *m, const double *x, double *M, double *w, int revert, int stop);
void orc_construct_M_hat(const orc_model *m, const double *x, double *M_hat, double *w);
void orc_transpose(double *M, unsigned n);

/* ---- LU log-det and in-place inverse (organic_dense_matrix.c:547-607) */
double orc_logdet_inv(const orc_model *m, double *M_hat, int *perm, int *sign);

/* ---- force (organic_dense_matrix.c:267-442, organic_hmc.c:86-134) */
void orc_construct_QM1P(const orc_model *m, const double *x, const double *M_hat, double *QM1P, double *w, int time);
void orc_construct_all_QM1P(const orc_model *m, const double *x, const double *M_hat, double *QM1P, double *w, int *perm);
void orc_refine_QM1P(const orc_model *m, const double *x, double *QM1P, double *w, int time);
void orc_force(const orc_model *m, const double *x, double *p_dot, const double *M_hat, double *QM1P, double *w, int *perm, double logdet);
size_t orc_force_workspace(const orc_model *m);   /* doubles needed for QM1P + w */

/* ---- bosons (organic_bosons.c) with an own FFT in FFTW conventions */
void orc_r2c_many(unsigned nt, unsigned howmany, const double *in, double *out_c); /* out_c: (nt/2+1)*howmany complex, interleaved */
void orc_c2r_many(unsigned nt, unsigned howmany, const double *in_c, double *out);  /* unnormalised */
void orc_sample_free_bosons(const orc_model *m, orc_mt *rng, double *x);
void orc_sample_fourier_momenta(const orc_model *m, orc_mt *rng, double *p);
double orc_energy_momenta(const orc_model *m, const double *p);
void orc_evolve_bosons(const orc_model *m, double *x, double *p, double h);
void orc_evolve_bosons_approx_step(const orc_model *m, double *x, const double *p, double h);
void orc_boson_force(const orc_model *m, const double *x, double *p_dot, int overwrite);
double orc_higher_matsubaras_norm(const orc_model *m, const double *x);
double orc_hamilton(const orc_model *m, const double *x, double logdet, double energyP);
void orc_update_x(const orc_model *m, double *x, double *p, double *p_dot, double h, double half);

/* ---- HMC (organic_hmc.c:167-353) */
typedef struct {
	orc_model *m;
	orc_mt rng;
	double *x, *p, *p_dot, *x_old, *M_hat, *M_hat_old, *QM1P, *R_0t, *w;
	int *perm;
	double logdet;
	unsigned steps;
	double traj_length;
	const double *contr_mat;
} orc_hmc;
int orc_hmc_init(orc_hmc *h, orc_model *m, uint64_t seed, unsigned steps, double traj_length, const double *contr_mat);
void orc_hmc_free(orc_hmc *h);
void orc_leap_frog(orc_hmc *h);
/* one trajectory; results (NUM_RES=12) and greens (nt*9) written when non-NULL */
int orc_trajectory(orc_hmc *h, double *results, double *greens, double *dH_out, FILE *res_out, FILE *corr_out);
void orc_cc_correlator(const orc_model *m, const double *x, const double *M_hat, double *R_t0, double *R_0t,
		double *w, double *greens, double logdet);
void orc_fprint_results(FILE *out, const double *res, unsigned n, unsigned r);
void orc_fprint_corr(FILE *out, const double *corr, const double *contr_mat, unsigned nn, unsigned nt);

/* ---- transient localisation (organic_trans_loc.c:25-114) */
void orc_construct_hamilton(const orc_model *m, const double *x, double *ham, double *w);
void orc_apply_current_to_v(const orc_model *m, const double *x, const double *v, double *w, unsigned k);
int orc_eigh(double *a, double *ev, unsigned n);  /* ascending eigenvalues; eigenvectors as contiguous chunks */
double orc_hist_trans_loc(const orc_model *m, const double *x, const double *contr_mat, double width, double *greens, unsigned nbins);
double orc_tl_width(const orc_model *m, unsigned nbins);

#ifdef __cplusplus
}
#endif
#endif
