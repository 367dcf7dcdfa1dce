/*
 * oracle_core.c -- CPU restatement: RNG, lattice, parameters, Taylor pairs, slice stencil,
 * chain, LU log-det and inverse.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#include "oracle.h"

/* ------------------------------------------------------------------------------------------
 * MT19937-64, restating simulation/mt19937-64.c:73-78 (seeding) and :108-143 (generation);
 * genrand64_real2 :158-161.  Published algorithm of Matsumoto & Nishimura (2004/9/29).
 */
#define ORC_NN 312
#define ORC_MM 156
void orc_mt_seed(orc_mt *m, uint64_t seed){
	m->s[0] = seed;
	for(int k = 1; k < ORC_NN; k++)
		m->s[k] = 6364136223846793005ULL * (m->s[k-1] ^ (m->s[k-1] >> 62)) + (uint64_t)k;
	m->i = ORC_NN;
}

uint64_t orc_mt_u64(orc_mt *m){
	static const uint64_t mag[2] = {0ULL, 0xB5026F5AA96619E9ULL};
	const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
	if(m->i >= ORC_NN){
		for(int k = 0; k < ORC_NN; k++){
			const uint64_t y = (m->s[k] & upper) | (m->s[(k+1) % ORC_NN] & lower);
			m->s[k] = m->s[(k + ORC_MM) % ORC_NN] ^ (y >> 1) ^ mag[y & 1ULL];
		}
		m->i = 0;
	}
	uint64_t y = m->s[m->i++];
	y ^= (y >> 29) & 0x5555555555555555ULL;
	y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
	y ^= (y << 37) & 0xFFF7EEE000000000ULL;
	y ^= (y >> 43);
	return y;
}

double orc_mt_real2(orc_mt *m){
	return (double)(orc_mt_u64(m) >> 11) * (1.0/9007199254740992.0);
}

/* Box-Muller pairs, organic_random.c:11-22: u1 = sqrt(-2 log U), u2 = 2 pi U', (cos, sin). */
void orc_random_vector(orc_mt *m, double *x, unsigned n){
	const double two_pi = 6.283185307179586;   /* M_2PI, organic_aux.h:10 */
	for(unsigned k = 0; k < n; k += 2){
		const double r = sqrt(-2 * log(orc_mt_real2(m)));
		const double phi = two_pi * orc_mt_real2(m);
		x[k] = r * cos(phi);
		x[(k+1) % n] = r * sin(phi);
	}
}

/* ------------------------------------------------------------------------------------------
 * Triangular parallelogram with periodic boundaries, organic_lattice.c:10-46.
 * nnt[i*6+d]: neighbour site, d = right, left, top right, bottom left, top left, bottom right.
 * nnt[(ns+i)*6+d]: link id (site*3 + direction) of the bond from i in direction d.
 */
unsigned *orc_construct_triangular(unsigned l1, unsigned l2, unsigned *ns_out, unsigned *nn_out){
	const unsigned ns = l1*l2;
	unsigned *nnt = malloc(12 * (size_t)ns * sizeof(unsigned));
	for(unsigned i = 0; i < ns; i++){
		const unsigned c1 = i % l1, c2 = i / l1;
		const unsigned r  = c2*l1 + (c1 + 1) % l1;
		const unsigned l  = c2*l1 + (c1 + l1 - 1) % l1;
		const unsigned up = (c2 + 1) % l2, dn = (c2 + l2 - 1) % l2;
		unsigned *n = nnt + 6*i, *k = nnt + 6*(ns + i);
		n[0] = r;
		n[1] = l;
		n[2] = up*l1 + c1;
		n[3] = dn*l1 + c1;
		n[4] = up*l1 + (c1 + l1 - 1) % l1;
		n[5] = dn*l1 + (c1 + 1) % l1;
		k[0] = 3*i;       k[1] = 3*n[1];
		k[2] = 3*i + 1;   k[3] = 3*n[3] + 1;
		k[4] = 3*i + 2;   k[5] = 3*n[5] + 2;
	}
	*ns_out = ns;
	*nn_out = 6;
	return nnt;
}

/* unit-cell projections, organic_lattice.c:48-60: rows = Cartesian x,y; columns = R, TR, TL */
double *orc_construct_contraction_mat(const double *a, const double *b, unsigned n){
	double *c = malloc(2 * n * sizeof(double));
	for(unsigned d = 0; d < 2; d++){
		c[d*n]     = a[d];
		c[d*n + 1] = .5 * (a[d] + b[d]);
		c[d*n + 2] = .5 * (b[d] - a[d]);
	}
	return c;
}

/* organic_lattice.c:62-106 */
double orc_contract_greens(const double *gr_all, const double *mat, unsigned nn, unsigned tau){
	const unsigned n = nn/2;
	const double *gr = gr_all + (size_t)n*n*tau;
	double c = 0;
	if(mat){
		/* contr_gr_mat: sum over Cartesian c of (sum_a mat[c,a] gr[a])^2, reading only the
		 * first row of the 3x3 block (organic_lattice.c:70-76) */
		for(unsigned d = 0; d < 2; d++){
			double s = 0;
			for(unsigned a = 0; a < n; a++) s += mat[d*n + a] * gr[a];
			c += s*s;
		}
	}else{
		/* contr_gr_triangle, organic_lattice.c:81-94 */
		for(unsigned a = 0; a < n; a++)
			for(unsigned b = 0; b <= a; b++) c += gr[a*n + b] + gr[b*n + a];
		c *= .5;
		c -= gr[n-1] + gr[(n-1)*n];
	}
	return c;
}

/* ------------------------------------------------------------------------------------------
 * Unit conversion, organic_main.c:99-116.  lambda is relative to J (kappa = lambda*t/x_avg).
 */
void orc_scale_params(double beta, const double *J, const double *lambda, double mu0, double omega0,
		unsigned nt, int static_disorder, double *t, double *kappa, double *mu, double *omega, double *x_avg_out){
	const double room_temperature = 0.025;   /* organic_main.c:17 */
	const double x_avg = static_disorder ? 1 : sqrt(.5 / tanh(.5 * omega0 / room_temperature));
	const double scale = static_disorder ? -beta : beta / nt;
	*mu = mu0 * scale;
	*omega = omega0 * scale;
	for(int k = 0; k < 3; k++){
		t[k] = J[k] * scale;
		kappa[k] = lambda[k] * t[k] / x_avg;
		*mu -= 2 * fabs(t[k]);
	}
	if(x_avg_out) *x_avg_out = x_avg;
}

/* ------------------------------------------------------------------------------------------
 * Taylor pairs: coefficient tables of simulation/taylor-coeff_pairs.h (even orders, whose
 * pair products equal the truncated exponential series) and the application order of
 * set_taylor_pair_order (organic_dense_matrix.c:133-159).
 */
static const double tab2[][2] = {{1, 1}};
static const double tab4[][2] = {{1.8294933739891903, 0.9404019519941707}, {0.1705066260108097, 1.5785318126884535}};
static const double tab6[][2] = {{2.2561549994606459, 0.8008468192975424}, {1.0805822170351833, 1.8245938269821296},
	{-0.3367372164958292, 1.5494473211106643}};
static const double tab8[][2] = {{2.5115385528762511, 0.6852625793459637}, {1.6810015288012661, 1.7481130845728929},
	{0.4249750401763088, 2.0321227226190966}, {-0.6175151218538260, 1.4284655522676216}};

int orc_set_taylor(orc_model *m, int which){
	const double (*tab)[2];
	int np;
	switch(which){
		case 2: tab = tab2; np = 1; break;
		case 4: tab = tab4; np = 2; break;
		case 6: tab = tab6; np = 3; break;
		case 8: tab = tab8; np = 4; break;
		default: return -1;
	}
	/* pair tuples alternate between the two ends of the table, every third from the low end */
	unsigned order[ORC_MAX_PAIRS];
	const unsigned tuples = np / 2;
	unsigned lo = 0, hi = tuples - 1;
	for(unsigned q = 0; q < tuples; q++){
		unsigned tup;
		if(q % 3 == 1) tup = lo++;
		else tup = hi--;
		order[2*q] = tup;
		order[2*q + 1] = np - 1 - tup;
	}
	if(np % 2) order[2*tuples] = tuples;
	m->np = np;
	for(int i = 0; i < np; i++){
		m->a_re[i] = .5 * tab[order[i]][0] / np;
		m->a_im[i] = .5 * tab[order[i]][1] / np;
	}
	return 0;
}

/* ------------------------------------------------------------------------------------------
 * w = (mu + K(x)) v, organic_dense_matrix.c:96-120 (scatter over forward bonds, in site order).
 */
void orc_apply_mu_K_to_v(const orc_model *m, const double *x, const double *v, double *w, double mu){
	const unsigned ns = m->ns;
	for(unsigned i = 0; i < ns; i++) w[i] = mu * v[i];
	for(unsigned i = 0; i < ns; i++){
		for(unsigned k = 0; k < 3; k++){
			const unsigned nb = m->nnt[i*6 + 2*k];
			const double f = m->t[k] - m->kappa[k] * x[i*3 + k];
			const double here = f * v[nb];
			const double there = f * v[i];
			w[i] += here;
			w[nb] += there;
		}
	}
}

/* M <- (mu+K) M column by column, organic_dense_matrix.c:123-131; w: ns*ns scratch */
void orc_apply_mu_K_to_M(const orc_model *m, const double *x, double *M, double *w){
	const unsigned ns = m->ns;
#pragma omp parallel for
	for(unsigned c = 0; c < ns; c++){
		double *col = M + (size_t)c*ns, *tmp = w + (size_t)c*ns;
		orc_apply_mu_K_to_v(m, x, col, tmp, m->mu);
		memcpy(col, tmp, ns * sizeof(double));
	}
}

/* one pair factor v <- (1 + 2Re(a) A + |a|^2 A^2) v, organic_dense_matrix.c:168-183; w: 2*ns */
static void orc_apply_pair(const orc_model *m, const double *x, double *v, double *w, int i){
	const unsigned ns = m->ns;
	const double a_lin = 2 * m->a_re[i];
	const double a_quad = pow(hypot(m->a_re[i], m->a_im[i]), 2);   /* pow(cabs(a), 2), organic_dense_matrix.c:169 */
	double *z = w + ns;
	orc_apply_mu_K_to_v(m, x, v, w, m->mu);
	if(m->a_im[i] != 0){
		orc_apply_mu_K_to_v(m, x, w, z, m->mu);
		for(unsigned s = 0; s < ns; s++) v[s] += a_lin * w[s] + a_quad * z[s];
	}else{
		for(unsigned s = 0; s < ns; s++) v[s] += a_lin * w[s];
	}
}

/* partial pair products, organic_dense_matrix.c:186-207: forward pairs [0, stop), or
 * (revert) pairs np-1 down to stop+1 */
void orc_apply_exp_mu_K_to_v(const orc_model *m, const double *x, double *v, double *w, int revert, int stop){
	if(revert){
		for(int i = m->np - 1; i > stop; i--) orc_apply_pair(m, x, v, w, i);
	}else{
		for(int i = 0; i < stop; i++) orc_apply_pair(m, x, v, w, i);
	}
}

/* organic_dense_matrix.c:209-216; w: 2*ns*ns */
void orc_apply_exp_mu_K_to_M(const orc_model *m, const double *x, double *M, double *w, int revert, int stop){
	const unsigned ns = m->ns;
#pragma omp parallel for
	for(unsigned c = 0; c < ns; c++)
		orc_apply_exp_mu_K_to_v(m, x, M + (size_t)c*ns, w + 2*(size_t)c*ns, revert, stop);
}

static void orc_identity(double *M, unsigned n){
	memset(M, 0, (size_t)n*n*sizeof(double));
	for(unsigned i = 0; i < n; i++) M[(size_t)i*(n+1)] = 1;
}

/* M_hat = 1 + B_0 B_1 ... B_{nt-1}, slices applied right to left, organic_dense_matrix.c:251-265 */
void orc_construct_M_hat(const orc_model *m, const double *x, double *M_hat, double *w){
	const unsigned ns = m->ns;
	orc_identity(M_hat, ns);
	for(int tau = (int)m->nt - 1; tau >= 0; tau--)
		orc_apply_exp_mu_K_to_M(m, x + (size_t)tau*ns*3, M_hat, w, 0, m->np);
	if(!m->mode.zero_filling && !m->mode.single_particle)
		for(unsigned i = 0; i < ns; i++) M_hat[(size_t)i*(ns+1)] += 1;
}

/* organic_dense_matrix.c:57-71 */
void orc_transpose(double *M, unsigned n){
#pragma omp parallel for
	for(unsigned i = 0; i < n; i++)
		for(unsigned j = i+1; j < n; j++){
			const double tmp = M[(size_t)i*n + j];
			M[(size_t)i*n + j] = M[(size_t)j*n + i];
			M[(size_t)j*n + i] = tmp;
		}
}

/* ------------------------------------------------------------------------------------------
 * logDetM_and_inv_M_hat, organic_dense_matrix.c:547-607: LAPACK dgetrf (partial pivoting,
 * first index of max |a|, column-major view of the row-major chunks) gives
 * log|det| = sum log|U_ii| and the sign from diag and pivots; the matrix is then replaced by
 * its inverse in place (dgetri).  The restatement solves A X = I with the same factors.
 * SINGLE_PARTICLE returns log(trace) and leaves the matrix untouched (:550-551).
 */
double orc_logdet_inv(const orc_model *m, double *A, int *perm, int *sign_out){
	const unsigned n = m->ns;
	if(m->mode.single_particle){
		double tr = 0;
		for(unsigned i = 0; i < n; i++) tr += A[(size_t)i*(n+1)];
		if(sign_out) *sign_out = 1;
		return log(tr);
	}
	if(m->mode.schur_form_inv) return NAN;   /* not restated (needs dgees) */
	double logdet = 0;
	int sign = 1;
#define AE(r,c) A[(size_t)(c)*n + (r)]
	for(unsigned j = 0; j < n; j++){
		unsigned p = j;
		double best = fabs(AE(j,j));
		for(unsigned r = j+1; r < n; r++) if(fabs(AE(r,j)) > best){ best = fabs(AE(r,j)); p = r; }
		perm[j] = (int)p + 1;                      /* LAPACK 1-based pivots */
		if(p != j) for(unsigned c = 0; c < n; c++){ const double t = AE(j,c); AE(j,c) = AE(p,c); AE(p,c) = t; }
		const double d = AE(j,j);
		if(d != 0){
			const double inv = 1/d;
			for(unsigned r = j+1; r < n; r++) AE(r,j) *= inv;
		}
#pragma omp parallel for if(n - j > 64)
		for(unsigned c = j+1; c < n; c++){
			const double u = AE(j,c);
			if(u != 0) for(unsigned r = j+1; r < n; r++) AE(r,c) -= AE(r,j) * u;
		}
	}
	for(unsigned i = 0; i < n; i++){
		const double d = AE(i,i);
		logdet += log(fabs(d));
		sign *= d > 0 ? 1 : -1;
		sign *= perm[i] == (int)i+1 ? 1 : -1;
	}
	/* inverse: X = U^{-1} L^{-1} P, column by column */
	double *X = malloc((size_t)n*n*sizeof(double));
#pragma omp parallel for
	for(unsigned c = 0; c < n; c++){
		double *xc = X + (size_t)c*n;
		for(unsigned r = 0; r < n; r++) xc[r] = r == c ? 1 : 0;
		for(unsigned j = 0; j < n; j++){          /* row swaps in pivot order */
			const unsigned p = (unsigned)perm[j] - 1;
			if(p != j){ const double t = xc[j]; xc[j] = xc[p]; xc[p] = t; }
		}
		for(unsigned j = 0; j < n; j++){          /* unit lower solve */
			const double v = xc[j];
			if(v != 0) for(unsigned r = j+1; r < n; r++) xc[r] -= AE(r,j) * v;
		}
		for(int j = (int)n - 1; j >= 0; j--){     /* upper solve */
			xc[j] /= AE(j,j);
			const double v = xc[j];
			if(v != 0) for(int r = 0; r < j; r++) xc[r] -= AE(r,j) * v;
		}
	}
#undef AE
	memcpy(A, X, (size_t)n*n*sizeof(double));
	free(X);
	if(sign_out) *sign_out = sign;
	return logdet;
}
