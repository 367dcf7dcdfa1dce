/*
 * oracle_tl.c -- CPU restatement of transient localisation (simulation/organic_trans_loc.c)
 * with an own symmetric eigensolver in place of LAPACK dsyev (Householder tridiagonalisation
 * followed by the implicit QL iteration with Wilkinson shifts, accumulating eigenvectors;
 * the published EISPACK tred2/tql2 method).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* construct_hamilton, organic_dense_matrix.c:218-224: ham = identity, then (mu+K) applied to
 * every column, with mu replaced by the caller's value (run with -mu, organic_trans_loc.c:39) */
void orc_construct_hamilton(const orc_model *m, const double *x, double *ham, double *w){
	const unsigned ns = m->ns;
	memset(ham, 0, (size_t)ns*ns*sizeof(double));
	for(unsigned i = 0; i < ns; i++) ham[(size_t)i*(ns+1)] = 1;
	for(unsigned c = 0; c < ns; c++){
		orc_apply_mu_K_to_v(m, x, ham + (size_t)c*ns, w, -m->mu);
		memcpy(ham + (size_t)c*ns, w, ns*sizeof(double));
	}
}

/* apply_current_to_v, organic_dense_matrix.c:226-249: antisymmetric hopping along direction k */
void orc_apply_current_to_v(const orc_model *m, const double *x, const double *v, double *w, unsigned k){
	const unsigned ns = m->ns;
	for(unsigned i = 0; i < ns; i++) w[i] = 0;
	for(unsigned i = 0; i < ns; i++){
		const unsigned nb = m->nnt[i*6 + 2*k];
		const double f = m->t[k] - m->kappa[k] * x[i*3 + k];
		w[i] += f * v[nb];
		w[nb] += -f * v[i];
	}
}

/* Symmetric eigen-decomposition of the n x n matrix a (either storage order; it is symmetric).
 * On return ev holds ascending eigenvalues and chunk i of a (a + i*n) the i-th eigenvector. */
int orc_eigh(double *a, double *d, unsigned n){
	double *e = calloc(n, sizeof(double));
	double *z = a;              /* z[r][c] = z[r*n + c], row-major working copy */
#define Z(r,c) z[(size_t)(r)*n + (c)]
	/* Householder reduction to tridiagonal form, accumulating the transformation */
	for(int i = (int)n - 1; i > 0; i--){
		const int l = i - 1;
		double h = 0, scale = 0;
		if(l > 0){
			for(int k = 0; k <= l; k++) scale += fabs(Z(i,k));
			if(scale == 0){
				e[i] = Z(i,l);
			}else{
				for(int k = 0; k <= l; k++){ Z(i,k) /= scale; h += Z(i,k)*Z(i,k); }
				double f = Z(i,l);
				double g = f >= 0 ? -sqrt(h) : sqrt(h);
				e[i] = scale*g;
				h -= f*g;
				Z(i,l) = f - g;
				f = 0;
				for(int j = 0; j <= l; j++){
					Z(j,i) = Z(i,j)/h;
					g = 0;
					for(int k = 0; k <= j; k++) g += Z(j,k)*Z(i,k);
					for(int k = j+1; k <= l; k++) g += Z(k,j)*Z(i,k);
					e[j] = g/h;
					f += e[j]*Z(i,j);
				}
				const double hh = f/(h+h);
				for(int j = 0; j <= l; j++){
					f = Z(i,j);
					e[j] = g = e[j] - hh*f;
					for(int k = 0; k <= j; k++) Z(j,k) -= f*e[k] + g*Z(i,k);
				}
			}
		}else{
			e[i] = Z(i,l);
		}
		d[i] = h;
	}
	d[0] = 0;
	e[0] = 0;
	for(unsigned i = 0; i < n; i++){
		const int l = (int)i - 1;
		if(d[i] != 0){
			for(int j = 0; j <= l; j++){
				double g = 0;
				for(int k = 0; k <= l; k++) g += Z(i,k)*Z(k,j);
				for(int k = 0; k <= l; k++) Z(k,j) -= g*Z(k,i);
			}
		}
		d[i] = Z(i,i);
		Z(i,i) = 1;
		for(int j = 0; j <= l; j++){ Z(j,i) = 0; Z(i,j) = 0; }
	}
	/* implicit QL with shifts on the tridiagonal (d, e) */
	for(unsigned i = 1; i < n; i++) e[i-1] = e[i];
	if(n) e[n-1] = 0;
	for(unsigned l = 0; l < n; l++){
		int iter = 0;
		unsigned mm;
		do{
			for(mm = l; mm + 1 < n; mm++){
				const double dd = fabs(d[mm]) + fabs(d[mm+1]);
				if(fabs(e[mm]) <= 2.220446049250313e-16*dd) break;
			}
			if(mm != l){
				if(iter++ == 60){ free(e); return -1; }
				double g = (d[l+1] - d[l]) / (2*e[l]);
				double r = hypot(g, 1.0);
				g = d[mm] - d[l] + e[l]/(g + (g >= 0 ? fabs(r) : -fabs(r)));
				double s = 1, c = 1, p = 0;
				int i;
				for(i = (int)mm - 1; i >= (int)l; i--){
					double f = s*e[i], b = c*e[i];
					e[i+1] = (r = hypot(f, g));
					if(r == 0){ d[i+1] -= p; e[mm] = 0; break; }
					s = f/r; c = g/r;
					g = d[i+1] - p;
					r = (d[i] - g)*s + 2*c*b;
					d[i+1] = g + (p = s*r);
					g = c*r - b;
					for(unsigned k = 0; k < n; k++){
						f = Z(k,i+1);
						Z(k,i+1) = s*Z(k,i) + c*f;
						Z(k,i) = c*Z(k,i) - s*f;
					}
				}
				if(r == 0 && i >= (int)l) continue;
				d[l] -= p; e[l] = g; e[mm] = 0;
			}
		}while(mm != l);
	}
	free(e);
	/* sort ascending; eigenvector j is column j of Z -> store as chunk j (transpose) */
	double *v = malloc((size_t)n*n*sizeof(double));
	unsigned *idx = malloc(n*sizeof(unsigned));
	for(unsigned i = 0; i < n; i++) idx[i] = i;
	for(unsigned i = 1; i < n; i++){
		const unsigned t = idx[i];
		unsigned j = i;
		while(j > 0 && d[idx[j-1]] > d[t]){ idx[j] = idx[j-1]; j--; }
		idx[j] = t;
	}
	double *ds = malloc(n*sizeof(double));
	for(unsigned j = 0; j < n; j++){
		ds[j] = d[idx[j]];
		for(unsigned r = 0; r < n; r++) v[(size_t)j*n + r] = Z(r, idx[j]);
	}
#undef Z
	memcpy(d, ds, n*sizeof(double));
	memcpy(a, v, (size_t)n*n*sizeof(double));
	free(ds); free(v); free(idx);
	return 0;
}

/* organic_trans_loc.c:99-100 */
double orc_tl_width(const orc_model *m, unsigned nbins){
	double width = -2*m->mu/nbins;
	for(unsigned i = 0; i < 3; i++) width += 8*fabs(m->kappa[i])/nbins;
	return width;
}

/* hist_trans_loc, organic_trans_loc.c:25-83, with the disorder vector x given (the caller
 * draws it with orc_random_vector, organic_trans_loc.c:38).  Returns results[s] = Z/N.
 * greens is accumulated (+=) as in the reference. */
double orc_hist_trans_loc(const orc_model *m, const double *x, const double *contr_mat, double width, double *greens, unsigned nbins){
	const unsigned ns = m->ns;
	double *ham = malloc((size_t)ns*ns*sizeof(double));
	double *w = malloc(3*(size_t)ns*sizeof(double));
	double *ev = malloc(ns*sizeof(double));
	orc_construct_hamilton(m, x, ham, w);
	orc_eigh(ham, ev, ns);
	double partition = 0;
	for(unsigned i = 0; i < ns; i++){
		for(unsigned k = 0; k < 3; k++) orc_apply_current_to_v(m, x, ham + (size_t)i*ns, w + (size_t)k*ns, k);
		for(unsigned j = 0; j < ns; j++){
			if(ev[j] < ev[i]) continue;
			const double boltzmann = exp(-ev[i]) + exp(-ev[j]);
			const double freq = ev[j] - ev[i];
			const unsigned bin = (unsigned)(freq / width + .5);
			double proj[3];
			for(unsigned k = 0; k < 3; k++){
				double s = 0;
				for(unsigned r = 0; r < ns; r++) s += ham[(size_t)j*ns + r] * w[(size_t)k*ns + r];
				proj[k] = s;
			}
			double I_nm = 0;   /* contr_gr_mat on the projection vector */
			for(unsigned c = 0; c < 2; c++){
				double s = 0;
				for(unsigned a = 0; a < 3; a++) s += contr_mat[c*3 + a] * proj[a];
				I_nm += s*s;
			}
			if(bin < nbins) greens[bin] += boltzmann * I_nm / (width * ns);
		}
		partition += exp(-ev[i]);
	}
	free(ham); free(w); free(ev);
	return partition / ns;
}
