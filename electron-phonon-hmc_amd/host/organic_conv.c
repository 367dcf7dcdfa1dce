/*
 * organic_conv.c -- post-processing of the transient-localisation histogram greens.bin, with the
 * command line and the two output tables of the reference's tool (simulation/organic_conv.c:116-164):
 *
 *     organic_conv greens.bin widths.txt sigma_w.txt [sigma_t.txt]
 *
 * greens.bin (organic_aux.c:82-90 format): n, the bin width h, then sigma_j at the frequencies
 * j h, j = 0..n-1, in units of 1/beta.  widths.txt (organic_aux.c:121-141 format): beta, nt and a
 * list of Lorentzian smearing widths in eV, each converted to units of 1/beta (g = width * beta).
 *
 * sigma_w.txt, one line per bin frequency w = i h (printed as w / beta), one column per width g:
 *     S_g(w) = (pi/2) tanhc(w/2) h sum_j sigma_j [L_g(w - j h) + L_g(w + j h)],
 *     L_g(x) = g / (pi (g^2 + x^2)),
 * and for g = 0 the unsmeared bin, S_0(w) = (pi/2) tanhc(w/2) sigma_i (S_0(0) = sigma_0).
 *
 * sigma_t.txt, one line per imaginary time t = i / nt, i = 0..nt/2:
 *     G_g(t) = h sum_j sigma_j K_g(t, j h),   K_g(t, w) = int dE L_g(E - w) k(t, E),
 * with the finite-temperature kernel k(t, E) = cosh((1/2 - t) E) / cosh(E/2) (K_0 = k).
 *
 * Quadrature of K_g.  The substitution E = w + g tan(theta) carries the Lorentzian measure onto
 * dtheta / pi over (-pi/2, pi/2):  K_g(t, w) = (1/pi) int k(t, w + g tan(theta)) dtheta.  The
 * integrand is smooth and lies in (0, 1]; for 0 < t < 1 it decays like exp(-min(t, 1-t) |E|), so it
 * vanishes with all its derivatives at both ends.  Adaptive Gauss-Kronrod (7-point Gauss embedded
 * in the 15-point Kronrod rule) bisects every panel until |K15 - G7| <= TOL * (panel length) / pi,
 * which bounds the error of K_g by about TOL = 1e-12 absolute (the Kronrod value of a panel is far
 * more accurate than the G7 comparison that bounds it).  At t = 0 (and t = 1) k = 1 and K_g = 1.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "organic.h"

#define PI 3.141592653589793
#define TOL 1e-12
#define MAX_DEPTH 40

/* Gauss-Kronrod 7-15 nodes on [-1, 1] (non-negative half; the odd indices are the Gauss nodes)
 * and their Kronrod and Gauss weights */
static const double gk_x[8] = {
	0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
	0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
	0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
	0.207784955007898467600689403773245, 0.0};
static const double gk_wk[8] = {
	0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
	0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
	0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
	0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk_wg[4] = {
	0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
	0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

/* k(t, E) = cosh((1/2 - t) E) / cosh(E / 2) for t in [0, 1/2], without overflow:
 * (e^{-t a} + e^{-(1-t) a}) / (1 + e^{-a}) with a = |E| */
static double kernel(double t, double E)
{
	const double a = fabs(E);
	return (exp(-t * a) + exp(-(1 - t) * a)) / (1 + exp(-a));
}

struct smear { double t, w, g; };

static double integrand(const struct smear *s, double theta)
{
	return kernel(s->t, s->w + s->g * tan(theta));
}

/* K15 and G7 sums of one panel [a, b] */
static void gk15(const struct smear *s, double a, double b, double *k15, double *g7)
{
	const double c = .5 * (a + b), r = .5 * (b - a);
	const double fc = integrand(s, c);
	double sk = gk_wk[7] * fc, sg = gk_wg[3] * fc;
	for(int i = 0; i < 7; i++){
		const double f = integrand(s, c - r * gk_x[i]) + integrand(s, c + r * gk_x[i]);
		sk += gk_wk[i] * f;
		if(i & 1) sg += gk_wg[i >> 1] * f;
	}
	*k15 = r * sk;
	*g7 = r * sg;
}

static double adapt(const struct smear *s, double a, double b, int depth)
{
	double k15, g7;
	gk15(s, a, b, &k15, &g7);
	/* a NaN estimate (NaN input) ends the bisection as a converged panel does */
	if(!(fabs(k15 - g7) > TOL * (b - a) / PI) || depth >= MAX_DEPTH) return k15;
	const double m = .5 * (a + b);
	return adapt(s, a, m, depth + 1) + adapt(s, m, b, depth + 1);
}

/* K_g(t, w) for t in [0, 1] */
static double smeared_kernel(double t, double w, double g)
{
	if(isnan(t)) return t;          /* t = 0/0 when widths.txt gives nt = 0 */
	if(t > .5) t = 1 - t;
	if(t == 0) return 1;
	if(g == 0) return kernel(t, w);
	const struct smear s = {t, w, g};
	/* the ends of (-pi/2, pi/2) are excluded by the open Kronrod nodes */
	return adapt(&s, -.5 * PI, .5 * PI, 0) / PI;
}

static double lorentzian(double x, double g){ return g / (PI * (g * g + x * x)); }

static double tanhc_half(double w){ return w != 0 ? tanh(.5 * w) / (.5 * w) : 1; }

static double smeared_sigma(double w, const double *sigma, unsigned n, double h, double g)
{
	if(g == 0){
		const unsigned bin = (unsigned)(w / h + .5);
		return bin == 0 ? sigma[0] : .5 * PI * tanhc_half(w) * sigma[bin];
	}
	double s = 0;
	for(unsigned j = 0; j < n; j++) s += sigma[j] * (lorentzian(w - j * h, g) + lorentzian(w + j * h, g));
	return .5 * PI * tanhc_half(w) * h * s;
}

static double correlator(double t, const double *sigma, unsigned n, double h, double g)
{
	double s = 0;
	for(unsigned j = 0; j < n; j++) s += sigma[j] * smeared_kernel(t, j * h, g);
	return h * s;
}

int main(int argc, char **argv)
{
	if(argc != 4 && argc != 5){
		printf("Error! 3 or 4 files needed: greens data, list of smearing widths, frequency output, opt. time output.\n");
		return 0;
	}
	unsigned n = 0, nw = 0, nt = 0;
	double h = 0, beta = 0;
	double *sigma = fread_greens(argv[1], &h, &n);
	double *widths = fscan_array(argv[2], &nw, &beta, &nt);
	if(!sigma || !widths){
		fprintf(stderr, "organic_conv: cannot read %s or %s\n", argv[1], argv[2]);
		return 1;
	}
	for(unsigned k = 0; k < nw; k++) widths[k] *= beta;

	FILE *fw = fopen(argv[3], "w");
	if(!fw){ fprintf(stderr, "organic_conv: cannot write %s\n", argv[3]); return 1; }
	for(unsigned i = 0; i < n; i++){
		const double w = i * h;
		fprintf(fw, "%g\t", w / beta);
		for(unsigned k = 0; k < nw; k++) fprintf(fw, "%g\t", smeared_sigma(w, sigma, n, h, widths[k]));
		fprintf(fw, "\n");
	}
	fclose(fw);

	if(argc == 5){
		FILE *ft = fopen(argv[4], "w");
		if(!ft){ fprintf(stderr, "organic_conv: cannot write %s\n", argv[4]); return 1; }
		for(unsigned i = 0; i <= nt / 2; i++){
			const double t = (double)i / nt;
			fprintf(ft, "%g\t", t);
			for(unsigned k = 0; k < nw; k++) fprintf(ft, "%g\t", correlator(t, sigma, n, h, widths[k]));
			fprintf(ft, "\n");
		}
		fclose(ft);
	}
	free(sigma);
	free(widths);
	return 0;
}
