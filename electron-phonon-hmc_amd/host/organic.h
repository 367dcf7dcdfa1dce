/*
 * organic.h -- the reference's host interface (simulation/organic_*.h), served by the MI355X
 * engine (libephmc).  Same function names, argument order and meaning; the hot-path functions
 * run on the GPU, the bookkeeping (lattice, RNG, flags, writers, statistics) on the host.
 *
 * The only type that differs is the FFT plan argument: the reference passes FFTW plans bound
 * to the field buffers x, x+dim, xc, xc+compl_dim (organic_hmc.c:317-322); here the same
 * argument position carries organic_plan handles, bound to the same buffers by
 * organic_plan_create, which own the GPU context of the run.
 */
#ifndef ORGANIC_GPU_H
#define ORGANIC_GPU_H

#include <stdio.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUM_RES 12            /* organic_aux.h:4 */
#define ROOM_TEMPERATURE 0.025

/* organic_flags.h:4-14 (same layout) */
typedef struct OrganicFlags{
	int schur_form_inv;
	int zero_filling;
	int static_disorder;
	int single_particle;
	int constant_force;
	int no_fourier_acc;
	int classical_fourier_acc;
	unsigned cfa_steps;
	double cfa_mass;
} organic_flags;

typedef struct organic_engine *organic_plan;

/* ---- organic_flags.c: flag tokens until EOF, unknown tokens ignored; additionally
 * "SEED <n>" fixes the random seed (the reference seeds with time(NULL)). */
void set_flags(FILE *in, organic_flags *mode);
int organic_seed(unsigned long *seed);          /* 1 and the seed when a SEED token was read */
void organic_set_seed(unsigned long seed);
const char *organic_state_file(void);           /* restart file of a STATE token, or NULL */
void organic_set_state_file(const char *path);

/* ---- organic_lattice.c */
unsigned *construct_triangular(unsigned l1, unsigned l2, unsigned *ns, unsigned *nn);
double *construct_contraction_mat(double *a, double *b, unsigned n);
double contr_gr_mat(double *gr, double *mat, unsigned n);
double contr_gr_triangle(double *gr, unsigned n);
double contract_greens(double *gr, double *mat, unsigned nn, unsigned tau);

/* ---- mt19937-64.c, organic_random.c */
void init_genrand64(unsigned long seed);
unsigned long genrand64_int64(void);
double genrand64_real2(void);
void random_vector(double *x, unsigned N);
void organic_rng_save(unsigned long long *words, int *pos);     /* 312 words */
void organic_rng_load(const unsigned long long *words, int pos);

/* ---- organic_aux.c */
unsigned ceil_half(unsigned n);
unsigned ceil_log2(unsigned n);
void fprint_corr(FILE *out, double *corr, double *contr_mat, unsigned nn, unsigned nt);
void fprint_results(FILE *out, double *res, unsigned n, unsigned r);
void fwrite_sigma(const char *name, double *greens, double width, unsigned n);
double *fread_greens(const char *in_name, double *h, unsigned *length);
void print_config(double *x, unsigned ns, unsigned nn, unsigned nt);
void fprint_config(const char *name, double *x, unsigned ns, unsigned nn, unsigned nt, const char *mode);
void fwrite_config(const char *name, double *x, unsigned n, const char *mode);
double *fscan_array(const char *in_name, unsigned *length, double *beta, unsigned *nt);
double average(double *x, unsigned n);
double average_sq(double *x, unsigned n);
void average_links(double *x, double *scale, unsigned ns, unsigned nn, unsigned nt, double *res);
double scalar_dot(double *x, double *y, unsigned d);

/* ---- engine binding (replaces the FFTW plan set-up of run_hmc, organic_hmc.c:317-322) */
organic_plan organic_plan_create(double *x, double complex *xc, double *t, double *kappa, double mu, double omega,
		unsigned flavors, unsigned l1, unsigned l2, unsigned nt, unsigned steps, double traj_length, organic_flags *mode);
void organic_plan_destroy(organic_plan plan);

/* ---- organic_dense_matrix.h.  GPU: the stencil and pair products, M_hat, the inverse, QM1P
 * and refine, the correlator, construct_hamilton.  Host: the small dense helpers (mat_mul,
 * mat_mul_tr, mat_mat_mul, construct_id, transpose, transpose_rect, trace), apply_current_to_v
 * (one vector) and phonon_correlator (sums over x). */
void mat_mul(double *m, double *x, double *y, unsigned n);
void mat_mul_tr(double *m, double *x, double *y, unsigned n);
void mat_mat_mul(double *m1, double *m2, double *y, unsigned n);
void construct_id(double *m, unsigned n);
void transpose(double *m, unsigned n);
void transpose_rect(double *x, double *y, unsigned n1, unsigned n2);
double trace(double *m, unsigned n);
void apply_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn);
void apply_mu_K_to_M(double *x, double *M, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn);
void set_taylor_pair_order(void);
void apply_K_plus_1_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn);
void apply_2nd_order_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, double complex a);
void apply_pair_expansion_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, int revert, int stop);
void apply_exp_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, int revert, int stop);
void apply_exp_mu_K_to_M(double *x, double *M, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, int revert, int stop);
void construct_hamilton(double *x, double *ham, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn);
void apply_current_to_v(double *x, double *v, double *w, double *t, double *kappa, unsigned *nnt, unsigned ns, unsigned nn, unsigned k);
void construct_M_hat(double *x, double *M_hat, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, organic_flags *mode);
void construct_QM1P(double *x, double *M_hat, double *QM1P, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, int time, organic_flags *mode);
void construct_all_QM1P(double *x, double *M_hat, double *QM1P, double *w, int *perm, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, organic_flags *mode);
void refine_QM1P(double *x, double *QM1P, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, int time);
void refine_QM1P_approx(double *x, double *QM1P, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, int time);
void cc_correlator(double *x, double *M_hat, double *R_t0, double *R_0t, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, double *greens, double logDetM, organic_flags *mode);
void phonon_correlator(double *x, unsigned ns, unsigned nn, unsigned nt, double *greens);
double logDetM_and_inv_M_hat(double *M_hat, double *w, int *p, unsigned ns, unsigned flavors, organic_flags *mode);

/* ---- organic_bosons.h.  GPU: the colouring, energies, flows and the boson force.  Host:
 * hot_start and higher_Matsubaras_norm (random stream, sums over x). */
void hot_start(double *x, double omega, unsigned ns, unsigned nn, unsigned nt);
double higher_Matsubaras_norm(double *x, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt);
void sample_free_bosons(double *x, double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft);
void sample_fourier_momenta(double *p, double complex *pc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft, organic_flags *mode);
double energy_fourier_momenta(double complex *pc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft);
void evolve_bosons(double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft, double h);
void evolve_bosons_approx_step(double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft, double h, organic_flags *mode);
void boson_force(double *x, double *p_dot, double omega, unsigned ns, unsigned nn, unsigned nt, int overwrite);

/* ---- organic_hmc.h (GPU; check_force is the host loop of finite differences over GPU log dets) */
double check_force(double *x, double *p, double *p_dot, double *M_hat, double *w, int *perm, double *t, double *kappa, double mu, double omega, unsigned flavors, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, organic_flags *mode);
double energy_momenta(double *p, double complex *pc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft, organic_flags *mode);
double hamilton(double *x, double *p, double logDetM, double energyP, double omega, unsigned flavors, unsigned ns, unsigned nn, unsigned nt);
void force(double *x, double *p_dot, double *M_hat, double *QM1P, double *w, int *perm, double *t, double *kappa, double mu, double omega, unsigned flavors, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, double logDetM, organic_flags *mode);
void update_x(double *x, double *p, double *p_dot, double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft, double h, double half, organic_flags *mode);
void leap_frog(double *x, double *p, double *p_dot, double complex *xc, double *M_hat, double *QM1P, double *w, int *perm, double *t, double *kappa, double mu, double omega, unsigned flavors, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, unsigned steps, double traj_length, const organic_plan *fft, organic_flags *mode);
short trajectory(double *x, double complex *xc, int *perm, double *t, double *kappa, double mu, double omega, unsigned flavors, double *contr_mat, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, unsigned steps, double traj_length, double *results, double *greens, int iteration, unsigned meas_freq, const organic_plan *fft, double *old, FILE *res_out, FILE *corr_out, organic_flags *mode);
unsigned run_hmc(double *greens, unsigned l1, unsigned l2, unsigned nt, double *t, double *kappa, double mu, double omega, unsigned flavors, double *contr_mat, unsigned steps, double traj_length, unsigned therm, unsigned meas, unsigned meas_freq, const char *res_name, const char *corr_name, organic_flags *mode);

/* ---- organic_trans_loc.h (GPU) */
void hist_trans_loc(double *x, double *t, double *kappa, double mu, double *contr_mat, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, double width, double *results, double *greens, FILE *res_out);
unsigned run_trans_loc(double *greens, unsigned l1, unsigned l2, unsigned nt, double *t, double *kappa, double mu, double *contr_mat, unsigned meas, const char *res_name, const char *corr_name);

#ifdef __cplusplus
}
#endif
#endif
