/* organic_io.c -- flags, result writers and statistics of the host (organic_flags.c,
 * organic_aux.c of the reference; same text and binary formats). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "organic.h"

static int seed_given = 0;
static unsigned long seed_value = 0;
static char state_path[1024];

const char *organic_state_file(void){ return state_path[0] ? state_path : NULL; }

void organic_set_state_file(const char *path)
{
	snprintf(state_path, sizeof(state_path), "%s", path ? path : "");
}

int organic_seed(unsigned long *seed)
{
	if(seed_given && seed) *seed = seed_value;
	return seed_given;
}

void organic_set_seed(unsigned long seed)
{
	seed_given = 1;
	seed_value = seed;
}

/* Flag tokens are read until EOF; unknown tokens are ignored (organic_flags.c:18-39).  In
 * addition "SEED <n>" fixes the seed (the reference seeds with time(NULL)) and "STATE <file>"
 * names a restart file (run_hmc resumes from it when it exists and writes it at the end); the
 * reference parser skips both like any unknown token, so such input files stay valid for it. */
void set_flags(FILE *in, organic_flags *mode)
{
	memset(mode, 0, sizeof(*mode));
	char tok[200];
	while(fscanf(in, "%199s\n", tok) != EOF){
		if(!strcmp(tok, "SCHUR_FORM_INV")) mode->schur_form_inv = 1;
		else if(!strcmp(tok, "ZERO_FILLING")) mode->zero_filling = 1;
		else if(!strcmp(tok, "STATIC_DISORDER")) mode->static_disorder = 1;
		else if(!strcmp(tok, "SINGLE_PARTICLE")) mode->single_particle = 1;
		else if(!strcmp(tok, "CONSTANT_FORCE")) mode->constant_force = 1;
		else if(!strcmp(tok, "NO_FOURIER_ACC")) mode->no_fourier_acc = 1;
		else if(!strcmp(tok, "CLASSICAL_FOURIER_ACC")){
			mode->classical_fourier_acc = 1;
			if(fscanf(in, "%u %lg\n", &mode->cfa_steps, &mode->cfa_mass) != 2){
				mode->cfa_steps = 1;
				mode->cfa_mass = 0;
			}
		}else if(!strcmp(tok, "SEED")){
			unsigned long s;
			if(fscanf(in, "%lu\n", &s) == 1) organic_set_seed(s);
		}else if(!strcmp(tok, "STATE")){
			char path[1024];
			if(fscanf(in, "%1023s\n", path) == 1) organic_set_state_file(path);
		}
	}
}

unsigned ceil_half(unsigned n){ return n % 2 ? n / 2 + 1 : n / 2; }

unsigned ceil_log2(unsigned n)
{
	unsigned l = 0;
	while(n > 1){ n = ceil_half(n); l++; }
	return l;
}

/* corr.csv: "tau<TAB>value" per slice, value = contract_greens(...) * nt^2 */
void fprint_corr(FILE *out, double *corr, double *contr_mat, unsigned nn, unsigned nt)
{
	if(!out) return;
	const unsigned nt2 = nt * nt;
	for(unsigned i = 0; i < nt; i++) fprintf(out, "%d\t%.15g\n", i, contract_greens(corr, contr_mat, nn, i) * nt2);
	fflush(out);
}

/* results.csv: r values "%.15g<TAB>" per line, n lines */
void fprint_results(FILE *out, double *res, unsigned n, unsigned r)
{
	if(!out) return;
	for(unsigned i = 0; i < n; i++){
		for(unsigned k = 0; k < r; k++, res++) fprintf(out, "%.15g\t", *res);
		fprintf(out, "\n");
	}
	fflush(out);
}

/* greens.bin: unsigned n, double width, double[n] */
void fwrite_sigma(const char *name, double *greens, double width, unsigned n)
{
	FILE *out = name ? fopen(name, "w") : NULL;
	if(!out){ fprintf(stderr, "fwrite_sigma: cannot open %s\n", name ? name : "(null)"); return; }
	fwrite(&n, sizeof(unsigned), 1, out);
	fwrite(&width, sizeof(double), 1, out);
	fwrite(greens, sizeof(double), n, out);
	fclose(out);
}

/* configuration dumps (organic_aux.c:92-119) */
void print_config(double *x, unsigned ns, unsigned nn, unsigned nt)
{
	for(unsigned tau = 0; tau < nt; tau++){
		for(unsigned i = 0; i < ns; i++){
			for(unsigned k = 0; k < nn / 2; k++, x++) printf("%g, ", *x);
			printf("\t");
		}
		printf("\n");
	}
}

void fprint_config(const char *name, double *x, unsigned ns, unsigned nn, unsigned nt, const char *mode)
{
	FILE *out = fopen(name, mode);
	if(!out) return;
	for(unsigned tau = 0; tau < nt; tau++)
		for(unsigned i = 0; i < ns; i++){
			for(unsigned k = 0; k < nn / 2; k++, x++) fprintf(out, "%.16g\t", *x);
			fprintf(out, "\n");
		}
	fprintf(out, "\n");
	fclose(out);
}

void fwrite_config(const char *name, double *x, unsigned n, const char *mode)
{
	FILE *out = fopen(name, mode);
	if(!out) return;
	fwrite(x, sizeof(double), n, out);
	fclose(out);
}

/* whitespace-separated file: beta, nt, then values until the end (organic_aux.c:121-141) */
double *fscan_array(const char *in_name, unsigned *length, double *beta, unsigned *nt)
{
	FILE *in = fopen(in_name, "r");
	if(!in) return NULL;
	if(beta && fscanf(in, "%lg ", beta) != 1) *beta = 0;
	if(nt && fscanf(in, "%u ", nt) != 1) *nt = 0;
	unsigned cap = 1024, k = 0;
	double *x = malloc(cap * sizeof(double));
	while(x && fscanf(in, "%lg ", x + k) > 0){
		if(++k == cap){
			cap *= 2;
			double *y = realloc(x, cap * sizeof(double));
			if(!y){ free(x); x = NULL; }
			x = y;
		}
	}
	*length = k;
	fclose(in);
	return x;
}

double *fread_greens(const char *in_name, double *h, unsigned *length)
{
	FILE *in = fopen(in_name, "r");
	if(!in) return NULL;
	if(fread(length, sizeof(unsigned), 1, in) != 1 || fread(h, sizeof(double), 1, in) != 1){ fclose(in); return NULL; }
	double *x = malloc(*length * sizeof(double));
	if(x && fread(x, sizeof(double), *length, in) != *length){ free(x); x = NULL; }
	fclose(in);
	return x;
}

double average(double *x, unsigned n)
{
	double s = 0;
	for(unsigned i = 0; i < n; i++) s += x[i];
	return s / n;
}

double average_sq(double *x, unsigned n)
{
	double s = 0;
	for(unsigned i = 0; i < n; i++) s += x[i] * x[i];
	return s / n;
}

/* per direction: sum over sites and slices / ns (organic_aux.c:199-210) */
void average_links(double *x, double *scale, unsigned ns, unsigned nn, unsigned nt, double *res)
{
	(void)scale;
	const size_t n = (size_t)ns * nt;
	for(unsigned k = 0; k < nn; k++) res[k] = 0;
	for(size_t i = 0; i < n; i++)
		for(unsigned k = 0; k < nn; k++) res[k] += x[i * nn + k];
	for(unsigned k = 0; k < nn; k++) res[k] /= ns;
}

double scalar_dot(double *x, double *y, unsigned d)
{
	double s = 0;
	for(unsigned i = 0; i < d; i++) s += x[i] * y[i];
	return s;
}
