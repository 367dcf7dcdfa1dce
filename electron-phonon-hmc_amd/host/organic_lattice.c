/* organic_lattice.c -- triangular lattice tables and unit-cell contractions
 * (simulation/organic_lattice.c:10-106 of the reference; same tables and results). */
#include <stdlib.h>
#include "organic.h"

/* Periodic l1 x l2 parallelogram.  nnt[i*6+d]: neighbour site in direction d = right, left,
 * top right, bottom left, top left, bottom right; nnt[(ns+i)*6+d]: link id (site*3 + k) of
 * that bond. */
unsigned *construct_triangular(unsigned l1, unsigned l2, unsigned *ns_out, unsigned *nn_out)
{
	const unsigned ns = l1 * l2;
	unsigned *t = malloc(12 * (size_t)ns * sizeof(unsigned));
	if(!t) return NULL;
	for(unsigned i = 0; i < ns; i++){
		const unsigned a = i % l1, b = i / l1, up = (b + 1) % l2, dn = (b + l2 - 1) % l2;
		unsigned *n = t + 6 * (size_t)i, *k = t + 6 * ((size_t)ns + i);
		n[0] = b * l1 + (a + 1) % l1;
		n[1] = b * l1 + (a + l1 - 1) % l1;
		n[2] = up * l1 + a;
		n[3] = dn * l1 + a;
		n[4] = up * l1 + (a + l1 - 1) % l1;
		n[5] = dn * l1 + (a + 1) % l1;
		k[0] = 3 * i;      k[1] = 3 * n[1];
		k[2] = 3 * i + 1;  k[3] = 3 * n[3] + 1;
		k[4] = 3 * i + 2;  k[5] = 3 * n[5] + 2;
	}
	*ns_out = ns;
	*nn_out = 6;
	return t;
}

/* Cartesian projections of the bond vectors R, TR, TL (rows x, y) */
double *construct_contraction_mat(double *a, double *b, unsigned n)
{
	double *m = malloc(2 * n * sizeof(double));
	if(!m) return NULL;
	for(unsigned d = 0; d < 2; d++){
		m[d * n] = a[d];
		m[d * n + 1] = .5 * (a[d] + b[d]);
		m[d * n + 2] = .5 * (b[d] - a[d]);
	}
	return m;
}

/* sum over Cartesian components of (sum_a mat[c,a] gr[a])^2; gr is read as a vector of n */
double contr_gr_mat(double *gr, double *mat, unsigned n)
{
	double out = 0;
	for(unsigned d = 0; d < 2; d++){
		double s = 0;
		for(unsigned a = 0; a < n; a++) s += mat[d * n + a] * gr[a];
		out += s * s;
	}
	return out;
}

/* symmetrised n x n block with the two corner terms removed */
double contr_gr_triangle(double *gr, unsigned n)
{
	double out = 0;
	for(unsigned a = 0; a < n; a++)
		for(unsigned b = 0; b <= a; b++) out += gr[a * n + b] + gr[b * n + a];
	out *= .5;
	out -= gr[n - 1] + gr[(n - 1) * n];
	return out;
}

double contract_greens(double *gr, double *mat, unsigned nn, unsigned tau)
{
	const unsigned n = nn / 2;
	double *g = gr + (size_t)n * n * tau;
	return mat ? contr_gr_mat(g, mat, n) : contr_gr_triangle(g, n);
}
