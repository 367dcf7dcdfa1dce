/* organic_rng.c -- MT19937-64 (Matsumoto & Nishimura, 2004/9/29 version; the generator of
 * simulation/mt19937-64.c) with one process-wide state as in the reference, and the Box-Muller
 * normals of organic_random.c:11-22. */
#include <math.h>
#include <stdint.h>
#include "organic.h"

#define MT_N 312
#define MT_M 156
static uint64_t mt_state[MT_N];
static int mt_pos = MT_N + 1;

void init_genrand64(unsigned long seed)
{
	mt_state[0] = (uint64_t)seed;
	for(int k = 1; k < MT_N; k++)
		mt_state[k] = 6364136223846793005ULL * (mt_state[k - 1] ^ (mt_state[k - 1] >> 62)) + (uint64_t)k;
	mt_pos = MT_N;
}

unsigned long genrand64_int64(void)
{
	const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL, matrix = 0xB5026F5AA96619E9ULL;
	if(mt_pos >= MT_N){
		if(mt_pos == MT_N + 1) init_genrand64(5489UL);
		for(int k = 0; k < MT_N; k++){
			const uint64_t y = (mt_state[k] & upper) | (mt_state[(k + 1) % MT_N] & lower);
			mt_state[k] = mt_state[(k + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1ULL) ? matrix : 0ULL);
		}
		mt_pos = 0;
	}
	uint64_t y = mt_state[mt_pos++];
	y ^= (y >> 29) & 0x5555555555555555ULL;
	y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
	y ^= (y << 37) & 0xFFF7EEE000000000ULL;
	y ^= y >> 43;
	return (unsigned long)y;
}

/* generator state for restart files: 312 words and the position */
void organic_rng_save(unsigned long long *words, int *pos)
{
	for(int k = 0; k < MT_N; k++) words[k] = mt_state[k];
	*pos = mt_pos;
}

void organic_rng_load(const unsigned long long *words, int pos)
{
	for(int k = 0; k < MT_N; k++) mt_state[k] = words[k];
	mt_pos = pos;
}

/* [0, 1) with 53-bit resolution */
double genrand64_real2(void)
{
	return (double)(genrand64_int64() >> 11) * (1.0 / 9007199254740992.0);
}

/* pairs (r cos phi, r sin phi), r = sqrt(-2 log U), phi = 2 pi U'; the last pair of an odd
 * count wraps onto element 0 */
void random_vector(double *x, unsigned n)
{
	const double two_pi = 6.283185307179586;
	for(unsigned k = 0; k < n; k += 2){
		const double r = sqrt(-2 * log(genrand64_real2()));
		const double phi = two_pi * genrand64_real2();
		x[k] = r * cos(phi);
		x[(k + 1) % n] = r * sin(phi);
	}
}
