/*
 * oracle_force.c -- CPU restatement of the fermion force: binary-splitting construction of
 * all QM1P_tau = B_{tau+1}..B_{nt-1} M_hat^{-1} B_0..B_{tau-1}, the exact derivative of the
 * Taylor-pair polynomial (refine) and the projection onto bonds.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* organic_aux.c:11-19 */
static unsigned ceil_half(unsigned n){ return n % 2 ? n/2 + 1 : n/2; }
static unsigned ceil_log2(unsigned n){ unsigned l = 0; while(n > 1){ n = ceil_half(n); l++; } return l; }

size_t orc_force_workspace(const orc_model *m){
	const size_t mat = (size_t)m->ns * m->ns;
	return (m->mode.constant_force ? 1 : m->nt) * mat + 5 * mat;
}

static void identity(double *M, unsigned n){
	memset(M, 0, (size_t)n*n*sizeof(double));
	for(unsigned i = 0; i < n; i++) M[(size_t)i*(n+1)] = 1;
}

/* refine_QM1P, organic_dense_matrix.c:383-442: QM1P <- sum over pairs j of
 *   a_lin_j Z_j + a_quad_j (A Z_j + Z_j A),  Z_j = Q_{<j} QM1P Q_{>j},
 * the exact derivative of the pair-product polynomial.  w: 5*mat (first 2*mat = stencil scratch). */
void orc_refine_QM1P(const orc_model *m, const double *x, double *QM1P, double *w, int time){
	const unsigned ns = m->ns;
	const size_t mat = (size_t)ns*ns;
	const double *xt = x + (size_t)time*ns*3;
	double *cp = w + 2*mat, *loc = cp + mat, *loc2 = loc + mat;
	memcpy(cp, QM1P, mat*sizeof(double));
	memset(QM1P, 0, mat*sizeof(double));
	for(int j = 0; j < m->np; j++){
		const int cplx = m->a_im[j] != 0;
		const double a_lin = 2*m->a_re[j];
		const double a_quad = pow(hypot(m->a_re[j], m->a_im[j]), 2);   /* pow(cabs(a), 2), organic_dense_matrix.c:169 */
		memcpy(loc, cp, mat*sizeof(double));
		orc_apply_exp_mu_K_to_M(m, xt, loc, w, 0, j);
		orc_transpose(loc, ns);
		orc_apply_exp_mu_K_to_M(m, xt, loc, w, 1, j);
		if(cplx){
			memcpy(loc2, loc, mat*sizeof(double));
			orc_apply_mu_K_to_M(m, xt, loc2, w);
			orc_transpose(loc2, ns);
		}
		orc_transpose(loc, ns);
		for(size_t e = 0; e < mat; e++) QM1P[e] += a_lin * loc[e];
		if(cplx){
			orc_apply_mu_K_to_M(m, xt, loc, w);
			for(size_t e = 0; e < mat; e++) QM1P[e] += a_quad * (loc[e] + loc2[e]);
		}
	}
}

/* construct_QM1P (non-Schur branch), organic_dense_matrix.c:267-309: QM1P for one slice. */
void orc_construct_QM1P(const orc_model *m, const double *x, const double *M_hat, double *QM1P, double *w, int time){
	const unsigned ns = m->ns, nt = m->nt;
	if(!m->mode.single_particle) memcpy(QM1P, M_hat, (size_t)ns*ns*sizeof(double));
	else identity(QM1P, ns);
	for(int tau = (int)nt - 1; tau > time; tau--)
		orc_apply_exp_mu_K_to_M(m, x + (size_t)tau*ns*3, QM1P, w, 0, m->np);
	orc_transpose(QM1P, ns);
	for(int tau = 0; tau < time; tau++)
		orc_apply_exp_mu_K_to_M(m, x + (size_t)tau*ns*3, QM1P, w, 0, m->np);
	orc_transpose(QM1P, ns);
	orc_refine_QM1P(m, x, QM1P, w, time);
}

/* construct_all_QM1P, organic_dense_matrix.c:311-381.  Binary splitting: in each of
 * ceil_log2(nt) rounds every populated slot i spawns slot i+chunk (a copy that receives
 * `chunk` more slices from the right) and receives `chunk` more slices from the left.
 * perm[i] counts left slices (taken from nt-1 downwards), perm[nt+i] right slices (from 0). */
void orc_construct_all_QM1P(const orc_model *m, const double *x, const double *M_hat, double *QM1P, double *w, int *perm){
	const unsigned ns = m->ns, nt = m->nt;
	const size_t mat = (size_t)ns*ns;
	unsigned remaining = nt - 1;
	if(!m->mode.single_particle) memcpy(QM1P, M_hat, mat*sizeof(double));
	else identity(QM1P, ns);
	for(unsigned i = 0; i < 2*nt; i++) perm[i] = 0;
	const unsigned rounds = ceil_log2(nt);
	for(unsigned r = 0; r < rounds; r++){
		const unsigned chunk = ceil_half(remaining);
		for(int i = (int)nt - 1; i >= 0; i--){
			if(!(i == 0 || perm[i] || perm[nt+i])) continue;
			double *left = QM1P + (size_t)i*mat;
			const unsigned tgt = (unsigned)i + chunk;
			if(tgt < nt && perm[tgt] == 0 && perm[nt+tgt] == 0){
				double *right = QM1P + (size_t)tgt*mat;
				memcpy(right, left, mat*sizeof(double));
				perm[tgt] = perm[i];
				perm[nt+tgt] = perm[nt+i];
				orc_transpose(right, ns);
				for(int tau = perm[nt+tgt]; tau < perm[nt+tgt] + (int)chunk; tau++)
					orc_apply_exp_mu_K_to_M(m, x + (size_t)tau*ns*3, right, w, 0, m->np);
				orc_transpose(right, ns);
				perm[nt+tgt] += (int)chunk;
			}
			const int top = (int)nt - 1 - perm[i];
			for(int tau = top; tau > top - (int)chunk; tau--)
				orc_apply_exp_mu_K_to_M(m, x + (size_t)tau*ns*3, left, w, 0, m->np);
			perm[i] += (int)chunk;
		}
		remaining -= chunk;
	}
}

/* force, organic_hmc.c:86-134: p_dot[tau,i,k] = -kappa_k * flavors * (Y[i,nb] + Y[nb,i]). */
void orc_force(const orc_model *m, const double *x, double *p_dot, const double *M_hat, double *QM1P, double *w, int *perm, double logdet){
	const unsigned ns = m->ns, nt = m->nt;
	const size_t mat = (size_t)ns*ns;
	if(m->mode.constant_force) orc_construct_QM1P(m, x, M_hat, QM1P, w, 0);
	else orc_construct_all_QM1P(m, x, M_hat, QM1P, w, perm);
	double *Q = QM1P;
	for(unsigned tau = 0; tau < nt; tau++){
		if(!m->mode.constant_force) orc_refine_QM1P(m, x, Q, w, (int)tau);
		for(unsigned i = 0; i < ns; i++){
			for(unsigned k = 0; k < 3; k++){
				const unsigned nb = m->nnt[i*6 + 2*k];
				double proj = Q[(size_t)i*ns + nb] + Q[(size_t)nb*ns + i];
				const size_t pos = ((size_t)tau*ns + i)*3 + k;
				p_dot[pos] = -m->kappa[k] * proj * m->flavors;
				if(m->mode.single_particle) p_dot[pos] *= exp(-logdet);
			}
		}
		if(!m->mode.constant_force) Q += mat;
	}
	if(m->mode.no_fourier_acc) orc_boson_force(m, x, p_dot, 0);
}
