/*
 * oracle_bosons.c -- CPU restatement of the Fourier-accelerated boson sector
 * (simulation/organic_bosons.c) with an own FFT that follows FFTW's conventions:
 * r2c  X[n] = sum_t x[t] exp(-2 pi i n t / nt),  n = 0..nt/2   (unnormalised)
 * c2r  x[t] = sum_{n=0}^{nt-1} X[n] exp(+2 pi i n t / nt) with X[nt-n] = conj X[n]
 * batched with howmany series of stride `howmany` and distance 1 (organic_hmc.c:317-322).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static const double PI = 3.141592653589793;

/* in-place complex FFT of length n (sign -1 forward, +1 backward), radix-2 when n is a power
 * of two, direct O(n^2) sum otherwise.  re/im interleaved. */
static void cfft(double *z, unsigned n, int sign, double *tmp){
	if((n & (n-1)) == 0){
		for(unsigned i = 1, j = 0; i < n; i++){
			unsigned bit = n >> 1;
			for(; j & bit; bit >>= 1) j ^= bit;
			j ^= bit;
			if(i < j){
				double t = z[2*i]; z[2*i] = z[2*j]; z[2*j] = t;
				t = z[2*i+1]; z[2*i+1] = z[2*j+1]; z[2*j+1] = t;
			}
		}
		for(unsigned len = 2; len <= n; len <<= 1){
			const double ang = sign * 2 * PI / len;
			for(unsigned i = 0; i < n; i += len){
				for(unsigned k = 0; k < len/2; k++){
					const double wr = cos(ang*k), wi = sin(ang*k);
					double *a = z + 2*(i+k), *b = z + 2*(i+k+len/2);
					const double br = b[0]*wr - b[1]*wi, bi = b[0]*wi + b[1]*wr;
					b[0] = a[0] - br; b[1] = a[1] - bi;
					a[0] += br; a[1] += bi;
				}
			}
		}
	}else{
		for(unsigned k = 0; k < n; k++){
			double sr = 0, si = 0;
			for(unsigned t = 0; t < n; t++){
				const double ang = sign * 2 * PI * (double)((unsigned long)k*t % n) / n;
				const double c = cos(ang), s = sin(ang);
				sr += z[2*t]*c - z[2*t+1]*s;
				si += z[2*t]*s + z[2*t+1]*c;
			}
			tmp[2*k] = sr; tmp[2*k+1] = si;
		}
		memcpy(z, tmp, 2*n*sizeof(double));
	}
}

void orc_r2c_many(unsigned nt, unsigned howmany, const double *in, double *out){
	const unsigned nc = nt/2 + 1;
#pragma omp parallel
	{
		double *z = malloc(4*nt*sizeof(double));
#pragma omp for
		for(unsigned l = 0; l < howmany; l++){
			for(unsigned t = 0; t < nt; t++){ z[2*t] = in[(size_t)t*howmany + l]; z[2*t+1] = 0; }
			cfft(z, nt, -1, z + 2*nt);
			for(unsigned n = 0; n < nc; n++){
				out[2*((size_t)n*howmany + l)] = z[2*n];
				out[2*((size_t)n*howmany + l) + 1] = z[2*n+1];
			}
		}
		free(z);
	}
}

void orc_c2r_many(unsigned nt, unsigned howmany, const double *in, double *out){
	const unsigned nc = nt/2 + 1;
#pragma omp parallel
	{
		double *z = malloc(4*nt*sizeof(double));
#pragma omp for
		for(unsigned l = 0; l < howmany; l++){
			for(unsigned n = 0; n < nc; n++){
				z[2*n] = in[2*((size_t)n*howmany + l)];
				z[2*n+1] = in[2*((size_t)n*howmany + l) + 1];
			}
			z[1] = 0;                         /* imaginary part of DC is ignored by c2r */
			if(nt % 2 == 0) z[2*(nt/2)+1] = 0; /* and of the Nyquist mode */
			for(unsigned n = nc; n < nt; n++){ z[2*n] = z[2*(nt-n)]; z[2*n+1] = -z[2*(nt-n)+1]; }
			cfft(z, nt, +1, z + 2*nt);
			for(unsigned t = 0; t < nt; t++) out[(size_t)t*howmany + l] = z[2*t];
		}
		free(z);
	}
}

static double freq2(double omega, unsigned n, unsigned nt){
	const double s = 2*sin(PI / nt * n);
	return omega + s*s / omega;   /* organic_bosons.c:64,114,137 */
}

static double fa_omega(const orc_model *m){
	/* classical Fourier acceleration uses a modified mass, organic_bosons.c:87,161 */
	return m->mode.classical_fourier_acc ? sqrt(m->mode.cfa_mass*m->mode.cfa_mass + m->omega*m->omega) : m->omega;
}

/* organic_bosons.c:55-74 */
void orc_sample_free_bosons(const orc_model *m, orc_mt *rng, double *x){
	const unsigned loc = m->ns*3, nt = m->nt, nc = nt/2 + 1;
	const size_t dim = (size_t)nt*loc;
	double *xc = malloc(2*(size_t)nc*loc*sizeof(double));
	orc_random_vector(rng, x, (unsigned)dim);
	orc_r2c_many(nt, loc, x, xc);
	for(unsigned n = 0; n < nc; n++){
		const double ampl = 1. / sqrt(freq2(m->omega, n, nt)) / nt;
		for(unsigned l = 0; l < 2*loc; l++) xc[2*(size_t)n*loc + l] *= ampl;
	}
	orc_c2r_many(nt, loc, xc, x);
	free(xc);
}

/* organic_bosons.c:76-100 */
void orc_sample_fourier_momenta(const orc_model *m, orc_mt *rng, double *p){
	const unsigned loc = m->ns*3, nt = m->nt, nc = nt/2 + 1;
	const size_t dim = (size_t)nt*loc;
	orc_random_vector(rng, p, (unsigned)dim);
	if(m->mode.no_fourier_acc) return;
	const double om = fa_omega(m);
	double *pc = malloc(2*(size_t)nc*loc*sizeof(double));
	orc_r2c_many(nt, loc, p, pc);
	for(unsigned n = 0; n < nc; n++){
		const double ampl = sqrt(freq2(om, n, nt)) / nt;
		for(unsigned l = 0; l < 2*loc; l++) pc[2*(size_t)n*loc + l] *= ampl;
	}
	orc_c2r_many(nt, loc, pc, p);
	free(pc);
}

/* energy_momenta, organic_hmc.c:56-65 -> energy_fourier_momenta, organic_bosons.c:102-124 */
double orc_energy_momenta(const orc_model *m, const double *p){
	const unsigned loc = m->ns*3, nt = m->nt, nc = nt/2 + 1, center = (nt+1)/2;
	const size_t dim = (size_t)nt*loc;
	if(m->mode.no_fourier_acc){
		double s = 0;
		for(size_t i = 0; i < dim; i++) s += p[i]*p[i];
		return .5 * dim * (s / dim);
	}
	const double om = fa_omega(m);
	double *pc = malloc(2*(size_t)nc*loc*sizeof(double));
	orc_r2c_many(nt, loc, p, pc);
	double en = 0;
	for(unsigned n = 0; n < nc; n++){
		double weight = .5 / nt / freq2(om, n, nt);
		if(n > 0 && n < center) weight *= 2;
		for(unsigned l = 0; l < loc; l++){
			const double re = pc[2*((size_t)n*loc + l)], im = pc[2*((size_t)n*loc + l) + 1];
			en += (re*re + im*im) * weight;
		}
	}
	free(pc);
	return en;
}

/* organic_bosons.c:126-150: exact rotation of each Matsubara mode */
void orc_evolve_bosons(const orc_model *m, double *x, double *p, double h){
	const unsigned loc = m->ns*3, nt = m->nt, nc = nt/2 + 1;
	double *xc = malloc(4*(size_t)nc*loc*sizeof(double)), *pc = xc + 2*(size_t)nc*loc;
	const double c = cos(h) / nt, s = sin(h) / nt;
	orc_r2c_many(nt, loc, x, xc);
	orc_r2c_many(nt, loc, p, pc);
	for(unsigned n = 0; n < nc; n++){
		const double f2 = freq2(m->omega, n, nt), sw = s*f2, s1w = s/f2;
		for(unsigned l = 0; l < 2*loc; l++){
			const size_t q = 2*(size_t)n*loc + l;
			const double x0 = xc[q], p0 = pc[q];
			xc[q] = c*x0 + s1w*p0;
			pc[q] = c*p0 - sw*x0;
		}
	}
	orc_c2r_many(nt, loc, xc, x);
	orc_c2r_many(nt, loc, pc, p);
	free(xc);
}

/* organic_bosons.c:152-177 (classical Fourier acceleration drift) */
void orc_evolve_bosons_approx_step(const orc_model *m, double *x, const double *p, double h){
	const unsigned loc = m->ns*3, nt = m->nt, nc = nt/2 + 1;
	double *xc = malloc(4*(size_t)nc*loc*sizeof(double)), *pc = xc + 2*(size_t)nc*loc;
	const double om = fa_omega(m), norm = 1. / nt;
	orc_r2c_many(nt, loc, x, xc);
	orc_r2c_many(nt, loc, p, pc);
	for(unsigned n = 0; n < nc; n++){
		const double inv_mass = h / freq2(om, n, nt);
		for(unsigned l = 0; l < 2*loc; l++){
			const size_t q = 2*(size_t)n*loc + l;
			xc[q] += inv_mass * pc[q];
			xc[q] *= norm;
		}
	}
	orc_c2r_many(nt, loc, xc, x);
	free(xc);
}

/* organic_bosons.c:179-194 */
void orc_boson_force(const orc_model *m, const double *x, double *p_dot, int overwrite){
	const unsigned loc = m->ns*3, nt = m->nt;
	if(overwrite) memset(p_dot, 0, (size_t)nt*loc*sizeof(double));
	for(unsigned t = 0; t < nt; t++){
		const size_t sh = (size_t)t*loc, shp = (size_t)((t+1) % nt)*loc, shm = (size_t)((t+nt-1) % nt)*loc;
		for(unsigned i = 0; i < loc; i++){
			p_dot[sh+i] -= m->omega * x[sh+i];
			p_dot[sh+i] += (x[shp+i] - 2*x[sh+i] + x[shm+i]) / m->omega;
		}
	}
}

/* organic_bosons.c:36-53 */
double orc_higher_matsubaras_norm(const orc_model *m, const double *x){
	const unsigned ns = m->ns, nt = m->nt;
	double xd2 = 0;
	for(unsigned i = 0; i < ns; i++)
		for(unsigned k = 0; k < 3; k++){
			double xm = 0;
			for(unsigned t = 0; t < nt; t++) xm += x[(size_t)t*ns*3 + i*3 + k];
			xm /= (double)nt;
			for(unsigned t = 0; t < nt; t++){
				const double d = x[(size_t)t*ns*3 + i*3 + k] - xm;
				xd2 += d*d;
			}
		}
	return xd2;
}

/* organic_hmc.c:67-84 (serial summation order) */
double orc_hamilton(const orc_model *m, const double *x, double logdet, double energyP){
	const unsigned ns = m->ns, nt = m->nt;
	double pot = 0, kin = 0;
	for(unsigned t = 0; t < nt; t++)
		for(unsigned i = 0; i < ns; i++){
			const size_t sh = ((size_t)t*ns + i)*3, shT = ((size_t)((t+1) % nt)*ns + i)*3;
			for(unsigned k = 0; k < 3; k++){
				pot += x[sh+k]*x[sh+k];
				const double d = x[shT+k] - x[sh+k];
				kin += d*d;
			}
		}
	return -logdet*m->flavors + energyP + (kin/m->omega + pot*m->omega)/2;
}

/* organic_hmc.c:136-165 */
void orc_update_x(const orc_model *m, double *x, double *p, double *p_dot, double h, double half){
	const size_t dim = (size_t)m->nt*m->ns*3;
	if(m->mode.no_fourier_acc){
		h *= fabs(half);
		for(size_t i = 0; i < dim; i++) x[i] += h*p[i];
	}else if(m->mode.classical_fourier_acc){
		h /= m->mode.cfa_steps;
		if(half != .5){
			orc_boson_force(m, x, p_dot, 1);
			for(size_t i = 0; i < dim; i++) p[i] += .5*h*p_dot[i];
			for(unsigned k = 0; k < m->mode.cfa_steps/2; k++){
				orc_evolve_bosons_approx_step(m, x, p, h);
				orc_boson_force(m, x, p_dot, 1);
				for(size_t i = 0; i < dim; i++) p[i] += h*p_dot[i];
			}
			orc_evolve_bosons_approx_step(m, x, p, .5*h);
		}
		if(half != -.5){
			orc_evolve_bosons_approx_step(m, x, p, .5*h);
			for(unsigned k = 0; k < m->mode.cfa_steps/2; k++){
				orc_boson_force(m, x, p_dot, 1);
				for(size_t i = 0; i < dim; i++) p[i] += h*p_dot[i];
				orc_evolve_bosons_approx_step(m, x, p, h);
			}
			orc_boson_force(m, x, p_dot, 1);
			for(size_t i = 0; i < dim; i++) p[i] += .5*h*p_dot[i];
		}
	}else orc_evolve_bosons(m, x, p, fabs(half)*h);
}
