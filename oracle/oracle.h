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
 * How each part is pinned (see tests/test_oracle_*.py):
 *   lattice tables, MT19937-64, Box-Muller normals, aux writers, flags: bit-exact against
 *     the reference's own sources compiled into oracle/_ref/libref.so;
 *   Taylor pairs: polynomial identity prod(pairs) == sum_{k<=4} z^k/k!;
 *   stencil, slice exponential, M_hat: dense numpy brute force on small lattices and the
 *     closed form for spatially uniform fields (Fourier diagonalisation);
 *   LU log-det / inverse: numpy slogdet / inv;
 *   force (construct_all_QM1P + refine + projection): central finite differences of the
 *     oracle's own log-det, the closed form for uniform fields, and the constant-force
 *     single-slice variant;
 *   FFT conventions: numpy rfft / irfft (unnormalised FFTW half-spectrum convention);
 *   boson flow / energies: exact invariants (kappa=0 conserves H, reversibility, Parseval);
 *   transient localisation: numpy eigh + brute-force binning.
 * Parity unpinned: SCHUR_FORM_INV (needs LAPACK dgees; not restated, returns an error).
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
