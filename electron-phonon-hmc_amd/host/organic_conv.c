/*
 * organic_conv.c -- post-processing of the transient-localisation histogram (greens.bin),
 * the counterpart of the reference's organic_conv (simulation/organic_conv.c:10-164):
 *
 *     organic_conv greens.bin widths.txt sigma_w.txt [sigma_t.txt]
 *
 * greens.bin is the binned conductivity written by run_trans_loc (unsigned n, bin width h,
 * n doubles).  widths.txt holds beta, nt and the list of smearing widths (organic_aux.c:121-141
 * format).  sigma_w.txt: per bin frequency (in units of 1/beta) the Lorentzian-smeared
 * conductivity for every width; sigma_t.txt: the imaginary-time correlator for t = 0 .. 1/2
 * obtained with the finite-temperature kernel (double-exponential quadrature).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "organic.h"

#define PI 3.141592653589793
#define HALF_PI 1.570796326794897

static double tanh_over_x(double x){ return x ? tanh(x) / x : 1; }

/* integrand of the time kernel at quadrature variable w for external frequency om */
static double kernel_integrand(double w, double t, double om, double width)
{
	const double eff = -(w * width + om);
	const double te = t * eff;
	return exp(te) / (1 + w * w) * (1 + exp(eff - 2 * te)) / (1 + exp(eff));
}

/* int_{-inf}^{inf} of the integrand, tanh-sinh substitution w = sinh(pi/2 sinh(x)), step .03,
 * summed until a term drops below DBL_EPSILON of the sum (at most 200 nodes per side) */
static double time_kernel_integral(double t, double om, double width)
{
	if(t > .5) t = 1 - t;
	if(t == 0) return PI;
	const double step = 0.03;
	double acc = HALF_PI * kernel_integrand(0, t, om, width);
	for(unsigned k = 1; k < 200; k++){
		const double x = k * step;
		const double s = HALF_PI * sinh(x), c = HALF_PI * cosh(x);
		const double w = sinh(s);
		const double f = (kernel_integrand(w, t, om, width) + kernel_integrand(w, t, -om, width)) * (c * cosh(s));
		if(f >= DBL_EPSILON * acc) acc += f;
		else break;
	}
	return step * acc;
}

/* Matsubara-sum form of the kernel with nt terms */
static double time_kernel_matsubara(double t, double om, double width, unsigned nt)
{
	const double complex z = width + I * om;
	double acc = creal(ccos(z * (t - .5)) / ccos(.5 * z));
	for(unsigned k = 0; k < nt; k++){
		const double wk = (2 * k + 1) * PI;
		const double complex d = I * wk - om;
		acc += creal(4 * width * sin(wk * t) / (width * width + d * d));
	}
	return acc;
}

static double time_kernel(double t, double om, double width, unsigned nt)
{
	if(nt) return time_kernel_matsubara(t, om, width, nt);
	if(width == 0) return kernel_integrand(0, t, om, 1);
	return time_kernel_integral(t, om, width) / PI;
}

static double lorentzian(double x, double width){ return width / (PI * (width * width + x * x)); }

/* smeared sigma(omega): sum_i sigma_i [L(omega - i h) + L(omega + i h)] h, with the tanh(w/2)/(w/2)
 * factor; width 0 returns the bin itself */
static double sigma_at(double om, const double *sigma, double width, double h, unsigned n)
{
	if(width == 0){
		const unsigned bin = (unsigned)(om / h + .5);
		if(bin == 0) return sigma[bin];
		return PI * .5 * tanh_over_x(.5 * om) * sigma[bin];
	}
	double acc = 0;
	for(unsigned i = 0; i < n; i++) acc += sigma[i] * (lorentzian(om - i * h, width) + lorentzian(om + i * h, width));
	return PI * .5 * tanh_over_x(.5 * om) * h * acc;
}

static double sigma_time(double t, const double *sigma, double width, double h, unsigned n, unsigned nt)
{
	double acc = 0;
	for(unsigned i = 0; i < n; i++) acc += sigma[i] * time_kernel(t, i * h, width, nt);
	return h * acc;
}

int main(int argc, char **argv)
{
	if(argc != 4 && argc != 5){
		printf("Error! 3 or 4 files needed: greens data, list of smearing widths, frequency output, opt. time output.\n");
		return 0;
	}
	unsigned n, nw, nt;
	double h, beta;
	double *sigma = fread_greens(argv[1], &h, &n);
	double *widths = fscan_array(argv[2], &nw, &beta, &nt);
	if(!sigma || !widths){
		fprintf(stderr, "organic_conv: cannot read %s\n", sigma ? argv[2] : argv[1]);
		return 1;
	}
	for(unsigned k = 0; k < nw; k++) widths[k] *= beta;
	FILE *out = fopen(argv[3], "w");
	for(unsigned i = 0; i < n; i++){
		const double om = i * h;
		fprintf(out, "%g\t", om / beta);
		for(unsigned k = 0; k < nw; k++) fprintf(out, "%g\t", sigma_at(om, sigma, widths[k], h, n));
		fprintf(out, "\n");
	}
	fclose(out);
	if(argc > 4){
		out = fopen(argv[4], "w");
		for(unsigned i = 0; i <= nt / 2; i++){
			const double t = (double)i / nt;
			fprintf(out, "%g\t", t);
			for(unsigned k = 0; k < nw; k++) fprintf(out, "%g\t", sigma_time(t, sigma, widths[k], h, n, 0));
			fprintf(out, "\n");
		}
		fclose(out);
	}
	free(sigma);
	free(widths);
	return 0;
}
