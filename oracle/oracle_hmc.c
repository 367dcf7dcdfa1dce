/*
 * oracle_hmc.c -- CPU restatement of the HMC driver (simulation/organic_hmc.c:167-353), the
 * current-current correlator (organic_dense_matrix.c:451-524) and the text writers
 * (organic_aux.c:51-70).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NUM_RES 12   /* organic_aux.h:4 */

int orc_hmc_init(orc_hmc *h, orc_model *m, uint64_t seed, unsigned steps, double traj_length, const double *contr_mat){
	const size_t dim = (size_t)m->nt*m->ns*3, mat = (size_t)m->ns*m->ns;
	memset(h, 0, sizeof(*h));
	h->m = m;
	h->steps = steps;
	h->traj_length = traj_length == 0 ? 1.570796326794897 : traj_length;   /* organic_main.c:114-116 */
	h->contr_mat = contr_mat;
	h->x = calloc(4*dim, sizeof(double));
	h->p = h->x + dim; h->p_dot = h->p + dim; h->x_old = h->p_dot + dim;
	h->M_hat = calloc(2*mat, sizeof(double));
	h->M_hat_old = h->M_hat + mat;
	/* workspace carve-up of organic_hmc.c:194-202: QM1P (nt*mat, also R_t0 of the
	 * correlator), R_0t (mat), w (5*mat) */
	h->QM1P = calloc(((size_t)m->nt + 6)*mat, sizeof(double));
	h->R_0t = h->QM1P + (size_t)m->nt*mat;
	h->w = h->R_0t + mat;
	h->perm = calloc(m->ns > 2*m->nt ? m->ns : 2*m->nt, sizeof(int));
	if(!h->x || !h->M_hat || !h->QM1P || !h->perm) return -1;
	orc_mt_seed(&h->rng, seed);
	/* run_hmc, organic_hmc.c:325-335 */
	orc_sample_free_bosons(m, &h->rng, h->x);
	orc_construct_M_hat(m, h->x, h->M_hat, h->w);
	h->logdet = 0;
	if(!m->mode.zero_filling) h->logdet = orc_logdet_inv(m, h->M_hat, h->perm, NULL);
	return 0;
}

void orc_hmc_free(orc_hmc *h){
	free(h->x); free(h->M_hat); free(h->QM1P); free(h->perm);
	memset(h, 0, sizeof(*h));
}

/* organic_hmc.c:167-190: x(h/2), then [force, p += h F, x(h)] with the last drift halved */
void orc_leap_frog(orc_hmc *h){
	const orc_model *m = h->m;
	const size_t dim = (size_t)m->nt*m->ns*3;
	const double step = h->steps ? h->traj_length/h->steps : 2*h->traj_length;
	double half = .5;
	orc_update_x(m, h->x, h->p, h->p_dot, step, half);
	half = 1;
	for(unsigned s = 1; s <= h->steps; s++){
		orc_construct_M_hat(m, h->x, h->M_hat, h->w);
		const double logdet = orc_logdet_inv(m, h->M_hat, h->perm, NULL);
		orc_force(m, h->x, h->p_dot, h->M_hat, h->QM1P, h->w, h->perm, logdet);
		for(size_t i = 0; i < dim; i++) h->p[i] += step*h->p_dot[i];
		if(s == h->steps) half = -.5;
		orc_update_x(m, h->x, h->p, h->p_dot, step, half);
	}
}

static double average(const double *x, size_t n){ double s = 0; for(size_t i = 0; i < n; i++) s += x[i]; return s/n; }
static double average_sq(const double *x, size_t n){ double s = 0; for(size_t i = 0; i < n; i++) s += x[i]*x[i]; return s/n; }
static double trace(const double *M, unsigned n){ double s = 0; for(unsigned i = 0; i < n; i++) s += M[(size_t)i*(n+1)]; return s; }

/* trajectory, organic_hmc.c:192-294.  Returns acceptance (0/1). */
int orc_trajectory(orc_hmc *h, double *results, double *greens, double *dH_out, FILE *res_out, FILE *corr_out){
	orc_model *m = h->m;
	const unsigned ns = m->ns, nt = m->nt;
	const size_t dim = (size_t)nt*ns*3, mat = (size_t)ns*ns;
	double logdet = 0, energy, energyP, boltzmann = 1, energy_old = 0;
	int acc = 1;
	if(!m->mode.zero_filling){
		memcpy(h->x_old, h->x, dim*sizeof(double));
		memcpy(h->M_hat_old, h->M_hat, mat*sizeof(double));
		orc_sample_fourier_momenta(m, &h->rng, h->p);
		energyP = orc_energy_momenta(m, h->p);
		energy_old = orc_hamilton(m, h->x, h->logdet, energyP);
		orc_leap_frog(h);
		orc_construct_M_hat(m, h->x, h->M_hat, h->w);
		logdet = orc_logdet_inv(m, h->M_hat, h->perm, NULL);
		energyP = orc_energy_momenta(m, h->p);
		energy = orc_hamilton(m, h->x, logdet, energyP);
		boltzmann = exp(energy_old - energy);
		if(dH_out) *dH_out = energy - energy_old;
		if(!(boltzmann > orc_mt_real2(&h->rng))){
			memcpy(h->x, h->x_old, dim*sizeof(double));
			memcpy(h->M_hat, h->M_hat_old, mat*sizeof(double));
			logdet = h->logdet;
			energy = energy_old;
			acc = 0;
		}
		h->logdet = logdet;
	}else{
		orc_sample_free_bosons(m, &h->rng, h->x);
		orc_construct_M_hat(m, h->x, h->M_hat, h->w);
		logdet = 0;
		energy = orc_hamilton(m, h->x, 0, 0);
		if(dH_out) *dH_out = 0;
	}
	if(greens){
		orc_cc_correlator(m, h->x, h->M_hat, h->QM1P, h->R_0t, h->w, greens, logdet);
		orc_fprint_corr(corr_out, greens, h->contr_mat, 6, nt);
	}
	if(results){
		results[0] = average(h->x, dim);
		results[1] = average_sq(h->x, dim) - 0.5;
		results[2] = energy;
		results[3] = -logdet;
		results[4] = orc_higher_matsubaras_norm(m, h->x);
		if(!m->mode.zero_filling && !m->mode.single_particle) results[5] = 1 - trace(h->M_hat, ns) / ns;
		else results[5] = trace(h->M_hat, ns) / ns;
		results[6] = acc;
		results[7] = boltzmann;
		/* average_links, organic_aux.c:199-210 */
		for(unsigned k = 0; k < 3; k++) results[8+k] = 0;
		for(size_t i = 0; i < (size_t)ns*nt; i++) for(unsigned k = 0; k < 3; k++) results[8+k] += h->x[i*3+k];
		for(unsigned k = 0; k < 3; k++) results[8+k] /= ns;
		if(greens) results[11] = orc_contract_greens(greens, h->contr_mat, 6, nt/2) * nt*nt / 3.141592653589793;
		orc_fprint_results(res_out, results, 1, NUM_RES);
	}
	return acc;
}

/* cc_correlator, organic_dense_matrix.c:451-524.  R_t0 needs nt*mat, R_0t mat, w 2*mat. */
void orc_cc_correlator(const orc_model *m, const double *x, const double *M_hat, double *R_t0, double *R_0t,
		double *w, double *greens, double logdet){
	const unsigned ns = m->ns, nt = m->nt;
	const size_t mat = (size_t)ns*ns;
	if(!m->mode.zero_filling && !m->mode.single_particle){
		memcpy(R_t0, M_hat, mat*sizeof(double));
		memcpy(R_0t, M_hat, mat*sizeof(double));
		orc_transpose(R_0t, ns);
	}else{
		memset(R_t0, 0, mat*sizeof(double)); memset(R_0t, 0, mat*sizeof(double));
		for(unsigned i = 0; i < ns; i++){ R_t0[(size_t)i*(ns+1)] = 1; R_0t[(size_t)i*(ns+1)] = 1; }
	}
	for(int tau = (int)nt - 1; tau >= 0; tau--){
		orc_apply_exp_mu_K_to_M(m, x + (size_t)tau*ns*3, R_t0, w, 0, m->np);
		if(tau) memcpy(R_t0 + (size_t)tau*mat, R_t0, mat*sizeof(double));
	}
	const double *Rt = R_t0;
	for(unsigned tau = 0; tau < nt; tau++, Rt += mat){
		const size_t shift = (size_t)tau*ns*3;
		for(unsigned a = 0; a < 3; a++){
			for(unsigned b = 0; b < 3; b++){
				double g = 0;
#pragma omp parallel for reduction(+:g)
				for(unsigned i = 0; i < ns; i++){
					for(unsigned j = 0; j < ns; j++){
						const unsigned ia = m->nnt[i*6 + 2*a], jb = m->nnt[j*6 + 2*b];
						const double chi_ia = m->t[a] - m->kappa[a] * x[i*3 + a];
						const double chi_jb = m->t[b] - m->kappa[b] * x[shift + j*3 + b];
						double corr = Rt[(size_t)j*ns + i] * R_0t[(size_t)jb*ns + ia];
						corr += Rt[(size_t)jb*ns + ia] * R_0t[(size_t)j*ns + i];
						corr -= Rt[(size_t)jb*ns + i] * R_0t[(size_t)j*ns + ia];
						corr -= Rt[(size_t)j*ns + ia] * R_0t[(size_t)jb*ns + i];
						g += chi_ia*chi_jb*corr;
					}
				}
				const size_t pos = (size_t)tau*9 + a*3 + b;
				greens[pos] = m->mode.single_particle ? g*exp(-logdet) : g/ns;
			}
		}
		if(tau < nt-1) orc_apply_exp_mu_K_to_M(m, x + shift, R_0t, w, 0, m->np);
	}
}

/* organic_aux.c:62-70 */
void orc_fprint_results(FILE *out, const double *res, unsigned n, unsigned r){
	if(!out) return;
	for(unsigned i = 0; i < n; i++){
		for(unsigned k = 0; k < r; k++) fprintf(out, "%.15g\t", res[(size_t)i*r + k]);
		fprintf(out, "\n");
	}
	fflush(out);
}

/* organic_aux.c:51-60 */
void orc_fprint_corr(FILE *out, const double *corr, const double *contr_mat, unsigned nn, unsigned nt){
	if(!out) return;
	for(unsigned i = 0; i < nt; i++)
		fprintf(out, "%d\t%.15g\n", i, orc_contract_greens(corr, contr_mat, nn, i) * (double)(nt*nt));
	fflush(out);
}
