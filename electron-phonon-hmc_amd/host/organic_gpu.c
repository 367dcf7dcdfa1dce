/*
 * organic_gpu.c -- the reference's hot-path functions (organic_dense_matrix.c, organic_hmc.c,
 * organic_bosons.c) with their call signatures, computed by the MI355X engine (libephmc).
 *
 * Functions taking host arrays copy their inputs to HBM, run the engine, and copy the outputs
 * back, so they are drop-in replacements one call at a time.  run_hmc keeps the chain state
 * (field, M_hat^{-1}, log det) resident on the GPU for the whole run and copies the field and
 * the inverse to the host only for measured trajectories.  The host keeps the random-number
 * stream (MT19937-64 and Box-Muller) in the reference's order, so a seeded run consumes
 * exactly the reference's draws: dim normals per trajectory followed by one uniform.
 *
 * Engine contexts are cached per model (lattice, slices, couplings); the FFT-plan argument of
 * the reference carries the context bound to the run's field buffers (organic_plan).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "ephmc.h"
#include "organic.h"

#define GPU(call)                                                                             \
	do {                                                                                      \
		int rc_ = (call);                                                                     \
		if(rc_ != EPHMC_OK){                                                                  \
			fprintf(stderr, "organic: %s failed (%d): %s\n", #call, rc_, ephmc_last_error()); \
			exit(EXIT_FAILURE);                                                               \
		}                                                                                     \
	} while(0)

struct organic_engine {
	ephmc_ctx *ctx;
	unsigned l1, l2, ns, nt, flavors, steps;
	double t[3], kappa[3], mu, omega, traj_length;
	int precision;
	organic_flags mode;
	size_t dim, mat;
	double *d_x, *d_p, *d_m, *d_v;          /* staging buffers for host-array calls */
	double *d_m2;                           /* second N x N staging buffer */
	double *d_q;                            /* nt N x N matrices of construct_all_QM1P (on first use) */
	double *x_host;                         /* buffers bound by organic_plan_create */
	double complex *xc_host;
	int chain_synced;                       /* device chain state equals the host field */
};

#define MAX_ENGINES 8
static struct organic_engine *engines[MAX_ENGINES];

/* Precision of the engine: FP64 drop-in unless EPHMC_PRECISION=mixed; partial products are
 * checkpointed every EPHMC_CHECKPOINT slices (default: the engine's choice, every slice while
 * they fit in 4 GiB). */
static int env_precision(void)
{
	const char *e = getenv("EPHMC_PRECISION");
	return (e && !strcmp(e, "mixed")) ? EPHMC_MIXED : EPHMC_FP64;
}

static unsigned env_checkpoint(void)
{
	const char *e = getenv("EPHMC_CHECKPOINT");
	return e ? (unsigned)atoi(e) : 0;
}

static const organic_flags no_flags;

static void check_flags(const organic_flags *mode, const char *who)
{
	if(mode && mode->schur_form_inv && mode->single_particle && mode->constant_force){
		fprintf(stderr, "organic: %s: SCHUR_FORM_INV with SINGLE_PARTICLE and CONSTANT_FORCE is not provided\n", who);
		exit(EXIT_FAILURE);
	}
}

/* lattice extents from a neighbour table of construct_triangular: the top-right neighbour of
 * site 0 is site l1 when l2 > 1 and site 0 when l2 == 1 */
static void lattice_extents(const unsigned *nnt, unsigned ns, unsigned *l1, unsigned *l2)
{
	*l1 = nnt[2] ? nnt[2] : ns;
	*l2 = ns / *l1;
	unsigned n2, nn;
	unsigned *ref = construct_triangular(*l1, *l2, &n2, &nn);
	if(!ref || n2 != ns || memcmp(ref, nnt, 12 * (size_t)ns * sizeof(unsigned))){
		fprintf(stderr, "organic: neighbour table is not a triangular parallelogram\n");
		exit(EXIT_FAILURE);
	}
	free(ref);
}

static int same(double a, double b){ return a == b; }

static void engine_release(struct organic_engine *e)
{
	ephmc_free(e->d_x); ephmc_free(e->d_p); ephmc_free(e->d_v); ephmc_free(e->d_m); ephmc_free(e->d_m2);
	ephmc_free(e->d_q);
	ephmc_destroy(e->ctx);
	free(e);
}

static struct organic_engine *engine_get(unsigned l1, unsigned l2, unsigned nt, const double *t, const double *kappa,
		double mu, double omega, unsigned flavors, unsigned steps, double traj_length, const organic_flags *mode)
{
	if(!mode) mode = &no_flags;
	check_flags(mode, "engine");
	for(int i = 0; i < MAX_ENGINES; i++){
		struct organic_engine *e = engines[i];
		if(e && e->l1 == l1 && e->l2 == l2 && e->nt == nt && same(e->mu, mu) && same(e->omega, omega)
			&& e->flavors == flavors && e->steps == steps && same(e->traj_length, traj_length)
			&& !memcmp(e->t, t, sizeof(e->t)) && !memcmp(e->kappa, kappa, sizeof(e->kappa))
			&& !memcmp(&e->mode, mode, sizeof(*mode)))
			return e;
	}
	struct organic_engine *e = calloc(1, sizeof(*e));
	ephmc_config cfg;
	memset(&cfg, 0, sizeof(cfg));
	cfg.l1 = l1; cfg.l2 = l2; cfg.nt = nt;
	memcpy(cfg.t, t, sizeof(cfg.t));
	memcpy(cfg.kappa, kappa, sizeof(cfg.kappa));
	cfg.mu = mu; cfg.omega = omega; cfg.flavors = flavors;
	cfg.steps = steps; cfg.traj_length = traj_length;
	cfg.taylor_order = 4;
	cfg.precision = env_precision();
	cfg.checkpoint = env_checkpoint();
	cfg.zero_filling = mode->zero_filling;
	cfg.single_particle = mode->single_particle;
	cfg.constant_force = mode->constant_force;
	cfg.no_fourier_acc = mode->no_fourier_acc;
	cfg.classical_fourier_acc = mode->classical_fourier_acc;
	cfg.cfa_steps = mode->cfa_steps;
	cfg.cfa_mass = mode->cfa_mass;
	cfg.schur_form_inv = mode->schur_form_inv;
	GPU(ephmc_create(&cfg, &e->ctx));
	/* EPHMC_AUTOTUNE=1: measured launch shapes (faster, but the choice follows timing noise and
	 * moves the force's rounding, so runs are no longer bit-reproducible) */
	const char *tune = getenv("EPHMC_AUTOTUNE");
	if(tune && atoi(tune) == 1) GPU(ephmc_autotune(e->ctx, NULL));
	e->l1 = l1; e->l2 = l2; e->nt = nt; e->ns = l1 * l2;
	memcpy(e->t, t, sizeof(e->t));
	memcpy(e->kappa, kappa, sizeof(e->kappa));
	e->mu = mu; e->omega = omega; e->flavors = flavors; e->steps = steps; e->traj_length = traj_length;
	e->precision = cfg.precision;
	e->mode = *mode;
	e->dim = (size_t)nt * 3 * e->ns;
	e->mat = (size_t)e->ns * e->ns;
	GPU(ephmc_malloc(e->dim * sizeof(double), (void **)&e->d_x));
	GPU(ephmc_malloc(e->dim * sizeof(double), (void **)&e->d_p));
	GPU(ephmc_malloc(e->dim * sizeof(double), (void **)&e->d_v));
	GPU(ephmc_malloc(e->mat * sizeof(double), (void **)&e->d_m));
	GPU(ephmc_malloc(e->mat * sizeof(double), (void **)&e->d_m2));
	int slot = 0;
	while(slot < MAX_ENGINES && engines[slot]) slot++;
	if(slot == MAX_ENGINES){
		/* evict the oldest */
		engine_release(engines[0]);
		memmove(engines, engines + 1, (MAX_ENGINES - 1) * sizeof(engines[0]));
		slot = MAX_ENGINES - 1;
	}
	engines[slot] = e;
	return e;
}

static struct organic_engine *engine_model(double *t, double *kappa, double mu, unsigned *nnt, unsigned ns, unsigned nt,
		const organic_flags *mode)
{
	unsigned l1, l2;
	lattice_extents(nnt, ns, &l1, &l2);
	const double omega = 1, traj = 0;
	return engine_get(l1, l2, nt, t, kappa, mu, omega, 1, 1, traj, mode);
}

static void up(struct organic_engine *e, double *dev, const double *host, size_t n)
{
	GPU(ephmc_memcpy_h2d(e->ctx, dev, host, n * sizeof(double)));
}

static void down(struct organic_engine *e, double *host, const double *dev, size_t n)
{
	GPU(ephmc_memcpy_d2h(e->ctx, host, dev, n * sizeof(double)));
}

/* ---- engine binding (the FFTW plans of run_hmc, organic_hmc.c:317-322) ------------------- */
organic_plan organic_plan_create(double *x, double complex *xc, double *t, double *kappa, double mu, double omega,
		unsigned flavors, unsigned l1, unsigned l2, unsigned nt, unsigned steps, double traj_length, organic_flags *mode)
{
	struct organic_engine *e = engine_get(l1, l2, nt, t, kappa, mu, omega, flavors, steps, traj_length, mode);
	e->x_host = x;
	e->xc_host = xc;
	e->chain_synced = 0;
	return e;
}

void organic_plan_destroy(organic_plan plan)
{
	if(!plan) return;
	for(int i = 0; i < MAX_ENGINES; i++)
		if(engines[i] == plan){
			engine_release(plan);
			engines[i] = NULL;
			return;
		}
}

/* ---- organic_dense_matrix.c ------------------------------------------------------------- */
void construct_id(double *m, unsigned n)
{
	memset(m, 0, (size_t)n * n * sizeof(double));
	for(unsigned i = 0; i < n; i++) m[(size_t)i * (n + 1)] = 1;
}

void transpose(double *m, unsigned n)
{
	for(unsigned i = 0; i < n; i++)
		for(unsigned j = i + 1; j < n; j++){
			const double tmp = m[(size_t)i * n + j];
			m[(size_t)i * n + j] = m[(size_t)j * n + i];
			m[(size_t)j * n + i] = tmp;
		}
}

double trace(double *m, unsigned n)
{
	double s = 0;
	for(unsigned i = 0; i < n; i++) s += m[(size_t)i * (n + 1)];
	return s;
}

/* organic_dense_matrix.c:209-216 */
void apply_exp_mu_K_to_M(double *x, double *M, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn, int revert, int stop)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, 1, NULL);
	up(e, e->d_x, x, 3 * (size_t)ns);
	up(e, e->d_m, M, e->mat);
	GPU(ephmc_apply_exp_M(e->ctx, e->d_x, e->d_m, revert, stop));
	down(e, M, e->d_m, e->mat);
}

/* organic_dense_matrix.c:251-265 */
void construct_M_hat(double *x, double *M_hat, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn, unsigned nt, organic_flags *mode)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, nt, mode);
	up(e, e->d_x, x, e->dim);
	GPU(ephmc_construct_M_hat(e->ctx, e->d_x, e->d_m));
	down(e, M_hat, e->d_m, e->mat);
}

/* organic_dense_matrix.c:547-607: log|det| and the inverse in place; p receives the (identity)
 * row interchanges of the factorisation in LAPACK's 1-based form */
double logDetM_and_inv_M_hat(double *M_hat, double *w, int *p, unsigned ns, unsigned flavors, organic_flags *mode)
{
	(void)w;
	check_flags(mode, "logDetM_and_inv_M_hat");
	if(mode && mode->single_particle) return log(trace(M_hat, ns));   /* M_hat unchanged */
	struct organic_engine *e = NULL;
	for(int i = 0; i < MAX_ENGINES && !e; i++)
		if(engines[i] && engines[i]->ns == ns && !engines[i]->mode.single_particle
			&& engines[i]->mode.schur_form_inv == (mode ? mode->schur_form_inv : 0)) e = engines[i];
	if(!e){
		const double z[3] = {0, 0, 0};
		organic_flags m = {0};
		if(mode) m.schur_form_inv = mode->schur_form_inv;
		e = engine_get(ns, 1, 1, z, z, 0, 1, flavors, 1, 0, &m);
	}
	up(e, e->d_m, M_hat, e->mat);
	double logdet;
	int sign;
	const int rc = ephmc_logdet_inv(e->ctx, e->d_m, &logdet, &sign);
	/* a singular M_hat is no error for the caller: log|det| = -inf, as dgetrf's zero U_ii gives
	 * in the reference; the inverse is then NaN (the Metropolis step rejects) */
	if(rc != EPHMC_OK && rc != EPHMC_ERR_NUMERIC) GPU(rc);
	down(e, M_hat, e->d_m, e->mat);
	if(p) GPU(ephmc_last_pivots(e->ctx, p));
	if(sign < 0 && flavors % 2) printf("Negative determinant!\n");
	return logdet;
}

/* organic_dense_matrix.c:451-524 */
void cc_correlator(double *x, double *M_hat, double *R_t0, double *R_0t, double *w, double *t, double *kappa, double mu,
		unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, double *greens, double logDetM, organic_flags *mode)
{
	(void)R_t0; (void)R_0t; (void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, nt, mode);
	up(e, e->d_x, x, e->dim);
	up(e, e->d_m, M_hat, e->mat);
	GPU(ephmc_cc_correlator(e->ctx, e->d_x, e->d_m, logDetM, greens));
}

/* ---- organic_hmc.c / organic_bosons.c ------------------------------------------------------ */
/* organic_bosons.c:36-53 */
double higher_Matsubaras_norm(double *x, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt)
{
	(void)nnt;
	const unsigned nn2 = nn / 2;
	double xd2 = 0;
	for(unsigned i = 0; i < ns; i++)
		for(unsigned k = 0; k < nn2; k++){
			double xm = 0;
			for(unsigned tau = 0; tau < nt; tau++) xm += x[(size_t)tau * ns * nn2 + i * nn2 + k];
			xm /= (double)nt;
			for(unsigned tau = 0; tau < nt; tau++){
				const double d = x[(size_t)tau * ns * nn2 + i * nn2 + k] - xm;
				xd2 += d * d;
			}
		}
	return xd2;
}

/* organic_hmc.c:86-134 (M_hat holds the inverse) */
void force(double *x, double *p_dot, double *M_hat, double *QM1P, double *w, int *perm, double *t, double *kappa, double mu,
		double omega, unsigned flavors, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, double logDetM, organic_flags *mode)
{
	(void)QM1P; (void)w; (void)perm; (void)nn;
	unsigned l1, l2;
	lattice_extents(nnt, ns, &l1, &l2);
	struct organic_engine *e = engine_get(l1, l2, nt, t, kappa, mu, omega, flavors, 1, 0, mode);
	up(e, e->d_x, x, e->dim);
	up(e, e->d_m, M_hat, e->mat);
	GPU(ephmc_force(e->ctx, e->d_x, e->d_m, logDetM, e->d_p));
	down(e, p_dot, e->d_p, e->dim);
}

/* organic_hmc.c:67-84 */
double hamilton(double *x, double *p, double logDetM, double energyP, double omega, unsigned flavors, unsigned ns,
		unsigned nn, unsigned nt)
{
	(void)p; (void)nn;
	struct organic_engine *e = NULL;
	for(int i = 0; i < MAX_ENGINES && !e; i++)
		if(engines[i] && engines[i]->ns == ns && engines[i]->nt == nt && same(engines[i]->omega, omega)
			&& engines[i]->flavors == flavors) e = engines[i];
	if(!e){
		const double z[3] = {0, 0, 0};
		e = engine_get(ns, 1, nt, z, z, 0, omega, flavors, 1, 0, NULL);
	}
	up(e, e->d_x, x, e->dim);
	double H;
	GPU(ephmc_hamilton(e->ctx, e->d_x, logDetM, energyP, &H));
	return H;
}

/* organic_bosons.c:55-74: the field buffer is the one bound to the plan */
void sample_free_bosons(double *x, double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt,
		const organic_plan *fft)
{
	(void)xc; (void)omega; (void)ns; (void)nn; (void)nt;
	struct organic_engine *e = fft[0];
	random_vector(x, (unsigned)e->dim);
	up(e, e->d_v, x, e->dim);
	GPU(ephmc_sample_free_bosons(e->ctx, e->d_v, e->d_x));
	down(e, x, e->d_x, e->dim);
	e->chain_synced = 0;
}

/* organic_bosons.c:76-100 */
void sample_fourier_momenta(double *p, double complex *pc, double omega, unsigned ns, unsigned nn, unsigned nt,
		const organic_plan *fft, organic_flags *mode)
{
	(void)pc; (void)omega; (void)ns; (void)nn; (void)nt;
	struct organic_engine *e = fft[0];
	random_vector(p, (unsigned)e->dim);
	up(e, e->d_v, p, e->dim);
	GPU(ephmc_sample_fourier_momenta(e->ctx, e->d_v, e->d_p));
	down(e, p, e->d_p, e->dim);
}

/* organic_bosons.c:102-124: reads the momenta of the plan-bound buffer (x + dim) */
double energy_fourier_momenta(double complex *pc, double omega, unsigned ns, unsigned nn, unsigned nt,
		const organic_plan *fft)
{
	(void)pc; (void)omega; (void)ns; (void)nn; (void)nt;
	struct organic_engine *e = fft[0];
	up(e, e->d_p, e->x_host + e->dim, e->dim);
	double en;
	GPU(ephmc_energy_momenta(e->ctx, e->d_p, &en));
	return en;
}

/* organic_hmc.c:56-65: p itself without Fourier acceleration, else the plan-bound momenta */
double energy_momenta(double *p, double complex *pc, double omega, unsigned ns, unsigned nn, unsigned nt,
		const organic_plan *fft, organic_flags *mode)
{
	if(!(mode && mode->no_fourier_acc)) return energy_fourier_momenta(pc, omega, ns, nn, nt, fft);
	struct organic_engine *e = fft[0];
	up(e, e->d_p, p, e->dim);
	double en;
	GPU(ephmc_energy_momenta(e->ctx, e->d_p, &en));
	return en;
}

/* organic_bosons.c:126-150 on the plan-bound field and momenta */
void evolve_bosons(double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt, const organic_plan *fft,
		double h)
{
	(void)xc; (void)omega; (void)ns; (void)nn; (void)nt;
	struct organic_engine *e = fft[0];
	up(e, e->d_x, e->x_host, e->dim);
	up(e, e->d_p, e->x_host + e->dim, e->dim);
	GPU(ephmc_evolve_bosons(e->ctx, e->d_x, e->d_p, h));
	down(e, e->x_host, e->d_x, e->dim);
	down(e, e->x_host + e->dim, e->d_p, e->dim);
	e->chain_synced = 0;
}

/* organic_hmc.c:136-165 on x and p (p_dot is the engine's own scratch) */
void update_x(double *x, double *p, double *p_dot, double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt,
		const organic_plan *fft, double h, double half, organic_flags *mode)
{
	(void)p_dot; (void)xc; (void)omega; (void)ns; (void)nn; (void)nt; (void)mode;
	struct organic_engine *e = fft[0];
	up(e, e->d_x, x, e->dim);
	up(e, e->d_p, p, e->dim);
	GPU(ephmc_update_x(e->ctx, e->d_x, e->d_p, h, half));
	down(e, x, e->d_x, e->dim);
	down(e, p, e->d_p, e->dim);
	e->chain_synced = 0;
}

/* organic_hmc.c:167-190: the whole leapfrog runs on the GPU; x and p are updated on return */
void leap_frog(double *x, double *p, double *p_dot, double complex *xc, double *M_hat, double *QM1P, double *w, int *perm,
		double *t, double *kappa, double mu, double omega, unsigned flavors, unsigned *nnt, unsigned ns, unsigned nn,
		unsigned nt, unsigned steps, double traj_length, const organic_plan *fft, organic_flags *mode)
{
	(void)p_dot; (void)xc; (void)M_hat; (void)QM1P; (void)w; (void)perm; (void)nnt; (void)nn; (void)fft;
	unsigned l1, l2;
	lattice_extents(nnt, ns, &l1, &l2);
	struct organic_engine *e = engine_get(l1, l2, nt, t, kappa, mu, omega, flavors, steps, traj_length, mode);
	up(e, e->d_x, x, e->dim);
	up(e, e->d_p, p, e->dim);
	GPU(ephmc_leap_frog(e->ctx, e->d_x, e->d_p));
	down(e, x, e->d_x, e->dim);
	down(e, p, e->d_p, e->dim);
}

/* Measured quantities of organic_hmc.c:256-291 from the host field and inverse. */
static void measure(struct organic_engine *e, double *x, double *M_hat, double energy, double logDetM, int acc,
		double boltzmann, double *contr_mat, unsigned *nnt, double *results, double *greens, FILE *res_out,
		FILE *corr_out, double *d_x, double *d_minv)
{
	const unsigned ns = e->ns, nt = e->nt;
	if(greens){
		GPU(ephmc_cc_correlator(e->ctx, d_x, d_minv, logDetM, greens));
		fprint_corr(corr_out, greens, contr_mat, 6, nt);
	}
	if(results){
		results[0] = average(x, (unsigned)e->dim);
		results[1] = average_sq(x, (unsigned)e->dim) - 0.5;
		results[2] = energy;
		results[3] = -logDetM;
		results[4] = higher_Matsubaras_norm(x, nnt, ns, 6, nt);
		if(!e->mode.zero_filling && !e->mode.single_particle) results[5] = 1 - trace(M_hat, ns) / ns;
		else results[5] = trace(M_hat, ns) / ns;     /* M_hat = B..B here */
		results[6] = acc;
		results[7] = boltzmann;
		average_links(x, e->kappa, ns, 3, nt, results + 8);
		if(greens) results[11] = contract_greens(greens, contr_mat, 6, nt / 2) * nt * nt / M_PI;
		fprint_results(res_out, results, 1, NUM_RES);
	}
}

/* organic_hmc.c:192-294.  The chain state lives on the GPU; x (the field, workspace start) and
 * the inverse at x + 4 dim are copied back after every call so that callers see the
 * reference's state. */
short trajectory(double *x, double complex *xc, int *perm, double *t, double *kappa, double mu, double omega,
		unsigned flavors, double *contr_mat, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, unsigned steps,
		double traj_length, double *results, double *greens, int iteration, unsigned meas_freq, const organic_plan *fft,
		double *old, FILE *res_out, FILE *corr_out, organic_flags *mode)
{
	(void)xc; (void)perm; (void)t; (void)kappa; (void)mu; (void)omega; (void)flavors; (void)nn; (void)steps;
	(void)traj_length; (void)mode;
	struct organic_engine *e = fft[0];
	const size_t dim = e->dim;
	double *p = x + dim, *M_hat = x + 4 * dim;
	if(!e->chain_synced){
		GPU(ephmc_chain_set_x(e->ctx, x));
		e->chain_synced = 1;
	}
	/* ZERO_FILLING draws a new free-boson field and no Metropolis uniform (organic_hmc.c:246-254) */
	random_vector(p, (unsigned)dim);
	const double u = e->mode.zero_filling ? 0.0 : genrand64_real2();
	ephmc_traj_record rec;
	GPU(ephmc_trajectory(e->ctx, p, u, &rec));
	GPU(ephmc_chain_get_x(e->ctx, x));
	GPU(ephmc_chain_get_minv(e->ctx, M_hat));
	old[0] = rec.logdet;
	if(iteration >= 0 && iteration % (int)meas_freq == 0){
		double *dx, *dp, *dm;
		GPU(ephmc_chain_buffers(e->ctx, &dx, &dp, &dm));
		measure(e, x, M_hat, rec.energy, rec.logdet, rec.accepted, rec.boltzmann, contr_mat, nnt, results, greens,
			res_out, corr_out, dx, dm);
	}
	return (short)rec.accepted;
}

/* Restart file of a STATE token: "EPHMCST1", dim (u64), the field x, the MT19937-64 state. */
static int state_load(const char *path, double *x, size_t dim)
{
	FILE *f = fopen(path, "rb");
	if(!f) return 0;
	char magic[8];
	unsigned long long d = 0, words[312];
	int pos = 0;
	const int ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, "EPHMCST1", 8) && fread(&d, sizeof(d), 1, f) == 1 && d == dim
		&& fread(x, sizeof(double), dim, f) == dim && fread(words, sizeof(words), 1, f) == 1 && fread(&pos, sizeof(pos), 1, f) == 1;
	fclose(f);
	if(!ok){
		fprintf(stderr, "organic: restart file %s does not belong to this lattice\n", path);
		exit(EXIT_FAILURE);
	}
	organic_rng_load(words, pos);
	return 1;
}

static void state_save(const char *path, const double *x, size_t dim)
{
	FILE *f = fopen(path, "wb");
	if(!f){
		fprintf(stderr, "organic: cannot write restart file %s\n", path);
		return;
	}
	unsigned long long d = dim, words[312];
	int pos;
	organic_rng_save(words, &pos);
	fwrite("EPHMCST1", 1, 8, f);
	fwrite(&d, sizeof(d), 1, f);
	fwrite(x, sizeof(double), dim, f);
	fwrite(words, sizeof(words), 1, f);
	fwrite(&pos, sizeof(pos), 1, f);
	fclose(f);
}

/* organic_hmc.c:296-353 with the chain resident on the GPU.  The seed is time(NULL) as in the
 * reference unless the input named one (SEED n); with STATE <file> the run resumes from the
 * file when it exists (field and random stream) and writes it at the end. */
unsigned run_hmc(double *greens, unsigned l1, unsigned l2, unsigned nt, double *t, double *kappa, double mu, double omega,
		unsigned flavors, double *contr_mat, unsigned steps, double traj_length, unsigned therm, unsigned meas,
		unsigned meas_freq, const char *res_name, const char *corr_name, organic_flags *mode)
{
	unsigned ns, nn;
	unsigned *nnt = construct_triangular(l1, l2, &ns, &nn);
	const unsigned greens_dim = nt * 9;
	const size_t dim = (size_t)nt * ns * 3, mat = (size_t)ns * ns;
	double *results = greens + (size_t)greens_dim * meas / meas_freq;
	double *x = malloc(dim * sizeof(double)), *normals = malloc(dim * sizeof(double));
	double *minv = malloc(mat * sizeof(double));
	FILE *res_out = res_name ? fopen(res_name, "a") : NULL;
	FILE *corr_out = corr_name ? fopen(corr_name, "a") : NULL;
	unsigned long seed;
	if(!organic_seed(&seed)) seed = (unsigned long)time(NULL);
	init_genrand64(seed);

	struct organic_engine *e = engine_get(l1, l2, nt, t, kappa, mu, omega, flavors, steps, traj_length, mode);
	const int zf = e->mode.zero_filling;
	const char *restart = organic_state_file();
	if(restart && state_load(restart, x, dim)){
		GPU(ephmc_chain_set_x(e->ctx, x));             /* resumes field and RNG stream */
	}else{
		random_vector(normals, (unsigned)dim);         /* sample_free_bosons' draw */
		GPU(ephmc_chain_init(e->ctx, normals));
	}
	double *dx, *dp, *dm;
	GPU(ephmc_chain_buffers(e->ctx, &dx, &dp, &dm));
	for(unsigned i = 0; i < therm + meas; i++){
		random_vector(normals, (unsigned)dim);
		const double u = zf ? 0.0 : genrand64_real2();
		const int measured = i >= therm && (i - therm) % meas_freq == 0;
		if(!measured){
			/* asynchronous: the next trajectory's normals are drawn while this one runs */
			GPU(ephmc_trajectory(e->ctx, normals, u, NULL));
			continue;
		}
		ephmc_traj_record rec;
		GPU(ephmc_trajectory(e->ctx, normals, u, &rec));
		GPU(ephmc_chain_get_x(e->ctx, x));
		GPU(ephmc_chain_get_minv(e->ctx, minv));
		const unsigned m = (i - therm) / meas_freq;
		measure(e, x, minv, rec.energy, rec.logdet, rec.accepted, rec.boltzmann, contr_mat, nnt,
			results + (size_t)m * NUM_RES, greens + (size_t)m * greens_dim, res_out, corr_out, dx, dm);
	}
	GPU(ephmc_synchronize(e->ctx));
	if(restart){
		GPU(ephmc_chain_get_x(e->ctx, x));
		state_save(restart, x, dim);
	}
	free(nnt); free(x); free(normals); free(minv);
	if(res_out) fclose(res_out);
	if(corr_out) fclose(corr_out);
	return meas / meas_freq * (greens_dim + NUM_RES);
}

/* ---- transient localisation (organic_trans_loc.c, organic_dense_matrix.c:218-249) --------- */
/* organic_dense_matrix.c:218-224: ham = (mu + K(x)) applied to the identity (one slice, no
 * exponential), on the GPU */
void construct_hamilton(double *x, double *ham, double *w, double *t, double *kappa, double mu, unsigned *nnt,
		unsigned ns, unsigned nn)
{
	(void)w;
	construct_id(ham, ns);
	apply_mu_K_to_M(x, ham, NULL, t, kappa, mu, nnt, ns, nn);
}

/* organic_dense_matrix.c:226-249: w = I_k v, the antisymmetric hopping along direction k (one
 * vector of N values: a host loop) */
void apply_current_to_v(double *x, double *v, double *w, double *t, double *kappa, unsigned *nnt, unsigned ns,
		unsigned nn, unsigned k)
{
	const unsigned nn2 = nn / 2;
	memset(w, 0, ns * sizeof(double));
	for(unsigned i = 0; i < ns; i++){
		const unsigned nb = nnt[i * nn + 2 * k];
		const double f = t[k] - kappa[k] * x[i * nn2 + k];
		w[i] += f * v[nb];
		w[nb] += -f * v[i];
	}
}

/* organic_trans_loc.c:87-114: meas independent disorder samples, each diagonalised and binned
 * on the GPU; the host draws the 3N normals of every sample in the reference's order. */
unsigned run_trans_loc(double *greens, unsigned l1, unsigned l2, unsigned nt, double *t, double *kappa, double mu,
		double *contr_mat, unsigned meas, const char *res_name, const char *corr_name)
{
	const unsigned nn2 = 3, ns = l1 * l2, dim = ns * nn2;
	double *results = greens + nt;
	FILE *res_out = res_name ? fopen(res_name, "a") : NULL;
	unsigned long seed;
	if(!organic_seed(&seed)) seed = (unsigned long)time(NULL);
	init_genrand64(seed);
	double width = -2 * mu / nt;
	for(unsigned k = 0; k < nn2; k++) width += 8 * fabs(kappa[k]) / nt;

	const double cm_default[6] = {0, 0, 0, 0, 0, 0};
	struct organic_engine *e = engine_get(l1, l2, 1, t, kappa, mu, 1, 1, 1, 0, NULL);
	/* samples are sent in batches of 2048 (the engine runs one wave per sample through the QL
	 * recurrence, so a batch puts two on every SIMD); each batch's normals are drawn in sample order */
	const unsigned batch = 2048;
	double *normals = malloc((size_t)batch * dim * sizeof(double));
	for(unsigned s0 = 0; s0 < meas; s0 += batch){
		const unsigned nb = meas - s0 < batch ? meas - s0 : batch;
		for(unsigned s = 0; s < nb; s++) random_vector(normals + (size_t)s * dim, dim);
		GPU(ephmc_tl_hist(e->ctx, normals, nb, contr_mat ? contr_mat : cm_default, width, nt, greens, results + s0));
		for(unsigned s = 0; s < nb; s++) fprint_results(res_out, results + s0 + s, 1, 1);
	}
	for(unsigned k = 0; k < nt; k++) greens[k] /= meas;
	fwrite_sigma(corr_name, greens, width, nt);
	free(normals);
	if(res_out) fclose(res_out);
	return meas + nt;
}

/* ---- the remaining functions of organic_dense_matrix.h, organic_hmc.h, organic_bosons.h and
 * organic_trans_loc.h ------------------------------------------------------------------------ */

/* Dense helpers of organic_dense_matrix.c:27-55 (not on the hot path; host loops). */
void mat_mul(double *m, double *x, double *y, unsigned n)
{
	for(unsigned i = 0; i < n; i++){
		double s = 0;
		for(unsigned k = 0; k < n; k++) s += m[(size_t)i * n + k] * x[k];
		y[i] = s;
	}
}

void mat_mul_tr(double *m, double *x, double *y, unsigned n)
{
	for(unsigned k = 0; k < n; k++) y[k] = 0;
	for(unsigned i = 0; i < n; i++)
		for(unsigned k = 0; k < n; k++) y[k] += m[(size_t)i * n + k] * x[i];
}

/* every chunk of m2 replaced by m1 times it; y: n scratch values */
void mat_mat_mul(double *m1, double *m2, double *y, unsigned n)
{
	for(unsigned c = 0; c < n; c++){
		double *col = m2 + (size_t)c * n;
		mat_mul(m1, col, y, n);
		memcpy(col, y, n * sizeof(double));
	}
}

/* y (n2 x n1) = x (n1 x n2) transposed */
void transpose_rect(double *x, double *y, unsigned n1, unsigned n2)
{
	for(unsigned i = 0; i < n1; i++)
		for(unsigned j = 0; j < n2; j++) y[(size_t)j * n1 + i] = x[(size_t)i * n2 + j];
}

/* The pair order is fixed per engine context when it is created (ephmc_create builds the
 * application order of organic_dense_matrix.c:133-159), so there is nothing left to set. */
void set_taylor_pair_order(void)
{
}

/* organic_dense_matrix.c:96-120: w = (mu + K(x)) v for the slice couplings x (GPU) */
void apply_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn)
{
	(void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, 1, NULL);
	up(e, e->d_x, x, 3 * (size_t)ns);
	up(e, e->d_v, v, ns);
	GPU(ephmc_apply_mu_K(e->ctx, e->d_x, e->d_v, e->d_p, 1));
	down(e, w, e->d_p, ns);
}

/* organic_dense_matrix.c:123-131: every chunk of M <- A chunk (GPU) */
void apply_mu_K_to_M(double *x, double *M, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, 1, NULL);
	up(e, e->d_x, x, 3 * (size_t)ns);
	up(e, e->d_m, M, e->mat);
	GPU(ephmc_apply_mu_K(e->ctx, e->d_x, e->d_m, e->d_m2, ns));
	down(e, M, e->d_m2, e->mat);
}

static void pairs_on_vector(double *x, double *v, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		int np, const double *al, const double *aq)
{
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, 1, NULL);
	up(e, e->d_x, x, 3 * (size_t)ns);
	up(e, e->d_v, v, ns);
	GPU(ephmc_apply_pairs(e->ctx, e->d_x, e->d_v, 1, np, al, aq));
	down(e, v, e->d_v, ns);
}

/* organic_dense_matrix.c:160-165: v <- (1 + A) v (GPU) */
void apply_K_plus_1_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn)
{
	(void)w; (void)nn;
	const double al = 1, aq = 0;
	pairs_on_vector(x, v, t, kappa, mu, nnt, ns, 1, &al, &aq);
}

/* organic_dense_matrix.c:168-183: v <- (1 + 2 Re(a) A + |a|^2 A^2) v, or (1 + 2 Re(a) A) v for a
 * real a (GPU) */
void apply_2nd_order_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt,
		unsigned ns, unsigned nn, double complex a)
{
	(void)w; (void)nn;
	const double al = 2 * creal(a), aq = cimag(a) != 0 ? pow(cabs(a), 2) : 0;
	pairs_on_vector(x, v, t, kappa, mu, nnt, ns, 1, &al, &aq);
}

/* organic_dense_matrix.c:186-194: the Taylor pairs [0, stop) or np-1 down to stop+1 on v (GPU) */
void apply_pair_expansion_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt,
		unsigned ns, unsigned nn, int revert, int stop)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, 1, NULL);
	up(e, e->d_x, x, 3 * (size_t)ns);
	up(e, e->d_v, v, ns);
	GPU(ephmc_apply_exp_cols(e->ctx, e->d_x, e->d_v, 1, revert, stop));
	down(e, v, e->d_v, ns);
}

/* organic_dense_matrix.c:197-207 (mu and K expanded together, as the reference's active branch) */
void apply_exp_mu_K_to_v(double *x, double *v, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn, int revert, int stop)
{
	apply_pair_expansion_mu_K_to_v(x, v, w, t, kappa, mu, nnt, ns, nn, revert, stop);
}

/* organic_dense_matrix.c:383-442 on the slice time of the field x (GPU) */
void refine_QM1P(double *x, double *QM1P, double *w, double *t, double *kappa, double mu, unsigned *nnt, unsigned ns,
		unsigned nn, int time)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, 1, NULL);
	up(e, e->d_x, x + (size_t)time * 3 * ns, 3 * (size_t)ns);
	up(e, e->d_m, QM1P, e->mat);
	GPU(ephmc_refine_QM1P(e->ctx, e->d_x, e->d_m));
	down(e, QM1P, e->d_m, e->mat);
}

/* organic_dense_matrix.c:444-449: the stencil in place of the polynomial's derivative */
void refine_QM1P_approx(double *x, double *QM1P, double *w, double *t, double *kappa, double mu, unsigned *nnt,
		unsigned ns, unsigned nn, int time)
{
	apply_mu_K_to_M(x + (size_t)time * 3 * ns, QM1P, w, t, kappa, mu, nnt, ns, nn);
}

/* organic_dense_matrix.c:267-309 (non-Schur branch; M_hat holds the inverse) (GPU) */
void construct_QM1P(double *x, double *M_hat, double *QM1P, double *w, double *t, double *kappa, double mu, unsigned *nnt,
		unsigned ns, unsigned nn, unsigned nt, int time, organic_flags *mode)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, nt, mode);
	up(e, e->d_x, x, e->dim);
	const int sp = mode && mode->single_particle;
	if(!sp) up(e, e->d_m, M_hat, e->mat);
	GPU(ephmc_construct_QM1P(e->ctx, e->d_x, sp ? NULL : e->d_m, e->d_m2, time));
	down(e, QM1P, e->d_m2, e->mat);
}

/* organic_dense_matrix.c:311-381: QM1P holds nt matrices; perm receives the counts of slices
 * taken from the left (perm[tau] = nt-1-tau) and from the right (perm[nt+tau] = tau) (GPU) */
void construct_all_QM1P(double *x, double *M_hat, double *QM1P, double *w, int *perm, double *t, double *kappa, double mu,
		unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, organic_flags *mode)
{
	(void)w; (void)nn;
	struct organic_engine *e = engine_model(t, kappa, mu, nnt, ns, nt, mode);
	if(!e->d_q) GPU(ephmc_malloc((size_t)nt * e->mat * sizeof(double), (void **)&e->d_q));
	up(e, e->d_x, x, e->dim);
	const int sp = mode && mode->single_particle;
	if(!sp) up(e, e->d_m, M_hat, e->mat);
	GPU(ephmc_construct_all_QM1P(e->ctx, e->d_x, sp ? NULL : e->d_m, e->d_q));
	down(e, QM1P, e->d_q, (size_t)nt * e->mat);
	if(perm)
		for(unsigned tau = 0; tau < nt; tau++){
			perm[tau] = (int)(nt - 1 - tau);
			perm[nt + tau] = (int)tau;
		}
}

/* organic_dense_matrix.c:526-545: equal-site phonon correlator <x_a(0) x_b(tau)> (host) */
void phonon_correlator(double *x, unsigned ns, unsigned nn, unsigned nt, double *greens)
{
	const unsigned nn2 = nn / 2;
	for(unsigned tau = 0; tau < nt; tau++)
		for(unsigned a = 0; a < nn2; a++)
			for(unsigned b = 0; b < nn2; b++){
				double g = 0;
				for(unsigned i = 0; i < ns; i++) g += x[(size_t)i * nn2 + a] * x[(size_t)tau * ns * nn2 + (size_t)i * nn2 + b];
				greens[(size_t)tau * nn2 * nn2 + a * nn2 + b] = g / ns;
			}
}

/* organic_hmc.c:20-54: relative error of p_dot against central differences of -flavors log det
 * (step DBL_EPSILON^(1/3)), over every component of x; each difference builds M_hat and its log
 * det on the GPU.  M_hat and w are scratch as in the reference. */
double check_force(double *x, double *p, double *p_dot, double *M_hat, double *w, int *perm, double *t, double *kappa,
		double mu, double omega, unsigned flavors, unsigned *nnt, unsigned ns, unsigned nn, unsigned nt, organic_flags *mode)
{
	(void)p; (void)omega;
	const double delta = cbrt(2.220446049250313e-16);
	const size_t dim = (size_t)nt * ns * nn / 2;
	double err = 0, norm = 0;
	for(size_t i = 0; i < dim; i++){
		const double keep = x[i];
		x[i] = keep + delta;
		construct_M_hat(x, M_hat, w, t, kappa, mu, nnt, ns, nn, nt, mode);
		const double lp = logDetM_and_inv_M_hat(M_hat, w, perm, ns, flavors, mode);
		x[i] = keep - delta;
		construct_M_hat(x, M_hat, w, t, kappa, mu, nnt, ns, nn, nt, mode);
		const double lm = logDetM_and_inv_M_hat(M_hat, w, perm, ns, flavors, mode);
		x[i] = keep;
		/* force = -dE/dx with E = -flavors log det */
		const double fd = flavors * (lp - lm) / (2 * delta);
		err += (p_dot[i] - fd) * (p_dot[i] - fd);
		norm += (fabs(p_dot[i]) + fabs(fd)) * (fabs(p_dot[i]) + fabs(fd));
	}
	return sqrt(err / norm);
}

/* organic_bosons.c:16-34: a field constant in time, each pair of link components drawn by one
 * Box-Muller step with the thermal spread sqrt(1/2 / tanh(nt omega / 2)) (host random stream) */
void hot_start(double *x, double omega, unsigned ns, unsigned nn, unsigned nt)
{
	const unsigned loc = ns * nn / 2;
	const double sd = sqrt(0.5 / tanh(0.5 * nt * omega));
	for(unsigned k = 0; k < loc; k += 2){
		const double r = sd * sqrt(-2 * log(genrand64_real2()));
		const double phi = 6.283185307179586 * genrand64_real2();
		for(unsigned tau = 0; tau < nt; tau++){
			x[(size_t)tau * loc + k] = r * cos(phi);
			x[(size_t)tau * loc + (k + 1) % loc] = r * sin(phi);
		}
	}
}

/* organic_bosons.c:152-177 on the plan-bound field x and momenta x + dim (GPU) */
void evolve_bosons_approx_step(double complex *xc, double omega, unsigned ns, unsigned nn, unsigned nt,
		const organic_plan *fft, double h, organic_flags *mode)
{
	(void)xc; (void)omega; (void)ns; (void)nn; (void)nt; (void)mode;
	struct organic_engine *e = fft[0];
	up(e, e->d_x, e->x_host, e->dim);
	up(e, e->d_p, e->x_host + e->dim, e->dim);
	GPU(ephmc_evolve_bosons_approx_step(e->ctx, e->d_x, e->d_p, h));
	down(e, e->x_host, e->d_x, e->dim);
	e->chain_synced = 0;
}

/* organic_bosons.c:179-194 (GPU) */
void boson_force(double *x, double *p_dot, double omega, unsigned ns, unsigned nn, unsigned nt, int overwrite)
{
	(void)nn;
	struct organic_engine *e = NULL;
	for(int i = 0; i < MAX_ENGINES && !e; i++)
		if(engines[i] && engines[i]->ns == ns && engines[i]->nt == nt && same(engines[i]->omega, omega)) e = engines[i];
	if(!e){
		const double z[3] = {0, 0, 0};
		e = engine_get(ns, 1, nt, z, z, 0, omega, 1, 1, 0, NULL);
	}
	up(e, e->d_x, x, e->dim);
	if(!overwrite) up(e, e->d_p, p_dot, e->dim);
	GPU(ephmc_boson_force(e->ctx, e->d_x, e->d_p, overwrite));
	down(e, p_dot, e->d_p, e->dim);
}

/* organic_trans_loc.c:25-83: one disorder sample drawn into x (3N normals, host random stream),
 * diagonalised and binned on the GPU; greens accumulates, results[0] = sum exp(-E_i) / N.  Pairs
 * whose bin index is nt or more are dropped (the reference writes them past the end of greens). */
void hist_trans_loc(double *x, double *t, double *kappa, double mu, double *contr_mat, unsigned *nnt, unsigned ns,
		unsigned nn, unsigned nt, double width, double *results, double *greens, FILE *res_out)
{
	unsigned l1, l2;
	lattice_extents(nnt, ns, &l1, &l2);
	const unsigned dim = ns * nn / 2;
	random_vector(x, dim);
	struct organic_engine *e = engine_get(l1, l2, 1, t, kappa, mu, 1, 1, 1, 0, NULL);
	GPU(ephmc_tl_hist(e->ctx, x, 1, contr_mat, width, nt, greens, results));
	fprint_results(res_out, results, 1, 1);
}
