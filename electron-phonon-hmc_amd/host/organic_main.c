/*
 * organic_main.c -- command-line driver with the reference's interface (organic_main.c):
 *     organic_main input.txt results.csv [corr.csv]
 * Reads the same input format, applies the same parameter scaling and runs HMC (or transient
 * localisation with STATIC_DISORDER) on the MI355X engine.  Extra input token: SEED <n>.
 * Environment: EPHMC_PRECISION=mixed selects FP32 molecular dynamics; EPHMC_CHECKPOINT=k the
 * interval of stored partial products; EPHMC_AUTOTUNE=1 measured launch shapes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "organic.h"

int main(int argc, char **argv)
{
	if(argc != 3 && argc != 4){
		printf("Error: One input and one or two output files needed!\n");
		return 0;
	}
	const char *corr_name = argc == 4 ? argv[3] : NULL;
	const unsigned nn2 = 3;
	FILE *in = fopen(argv[1], "r");
	if(!in){
		printf("Error: cannot open %s!\n", argv[1]);
		return 0;
	}
	unsigned l1, l2, nt, steps, therm, meas, meas_freq, flavors;
	double beta, t[3], kappa[3], mu, omega, traj_length, a[2], b[2];
	if(fscanf(in, "%u %u\n%u\n", &l1, &l2, &nt) != 3){ printf("Error: Need L1 L2, Nt!\n"); return 0; }
	if(fscanf(in, "%lg\n", &beta) != 1){ printf("Error: Need beta!\n"); return 0; }
	if(fscanf(in, "%lg %lg %lg\n%lg %lg %lg\n", t, t + 1, t + 2, kappa, kappa + 1, kappa + 2) != 6){
		printf("Error: Need 3 x t, 3 x kappa!\n");
		return 0;
	}
	if(fscanf(in, "%lg\n%lg\n", &mu, &omega) != 2){ printf("Error: Need mu0, omega!\n"); return 0; }
	if(fscanf(in, "%u %lg\n", &steps, &traj_length) != 2){ printf("Error: Need steps, traj-length!\n"); return 0; }
	if(fscanf(in, "%u\n%u %u\n", &therm, &meas, &meas_freq) != 3){
		printf("Error: Need therm, meas, meas-freq!\n");
		return 0;
	}
	double *contr_mat = NULL;
	if(fscanf(in, "%lg %lg\n%lg %lg\n", a, a + 1, b, b + 1) == 4) contr_mat = construct_contraction_mat(a, b, nn2);
	if(fscanf(in, "%u\n", &flavors) != 1) flavors = 1;
	organic_flags mode[1];
	set_flags(in, mode);
	fclose(in);

	/* organic_main.c:99-116: couplings in lattice units */
	const double x_avg = mode->static_disorder ? 1 : sqrt(.5 / tanh(.5 * omega / ROOM_TEMPERATURE));
	if(mode->static_disorder) beta = -beta;
	else beta /= nt;
	mu *= beta;
	omega *= beta;
	if(mode->classical_fourier_acc) mode->cfa_mass *= beta;
	for(unsigned k = 0; k < nn2; k++){
		t[k] *= beta;
		kappa[k] *= t[k] / x_avg;
		mu -= 2 * fabs(t[k]);
	}
	if(traj_length == 0) traj_length = 1.570796326794897;   /* organic_aux.h:10 */

	unsigned greens_dim, res_dim, length;
	if(mode->static_disorder){
		greens_dim = nt;                 /* number of conductivity bins */
		res_dim = meas + greens_dim;
	}else{
		greens_dim = nt * nn2 * nn2;
		res_dim = meas / meas_freq * (greens_dim + NUM_RES);
	}
	double *greens = calloc(res_dim, sizeof(double));
	if(mode->static_disorder)
		length = run_trans_loc(greens, l1, l2, nt, t, kappa, mu, contr_mat, meas, argv[2], corr_name);
	else
		length = run_hmc(greens, l1, l2, nt, t, kappa, mu, omega, flavors, contr_mat, steps, traj_length, therm, meas,
			meas_freq, argv[2], corr_name, mode);
	if(length != res_dim) printf("ERROR: Result vector expected of length %d, is %d!\n", res_dim, length);
	free(contr_mat);
	free(greens);
	return 0;
}
