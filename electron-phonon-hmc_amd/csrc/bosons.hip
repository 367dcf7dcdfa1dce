// bosons.hip -- Fourier-accelerated boson sector (organic_bosons.c) as batched FFT kernels,
// plus the deterministic reductions of the Hamiltonian and the Metropolis step.
//
// Every link l carries a real series over tau (stride 3N = loc, organic_hmc.c:317-322).  A
// workgroup loads LW complex series into LDS (z[t][l], links contiguous, coalesced rows),
// transforms them (radix-2 decimation in frequency: natural order in, bit-reversed modes out),
// applies the per-mode operation, and transforms back (decimation in time: bit-reversed in,
// natural out), so no permutation pass is needed.  Non-power-of-two nt uses direct sums.
// Two real series share one complex transform:
//   evolve:  z = x + i p of one link; with X_n = (Z_n + conj Z_-n)/2, P_n = (Z_n - conj Z_-n)/2i
//            the exact flow X' = cos h X + sin h P/f_n, P' = cos h P - sin h f_n X per mode
//            (evolve_bosons, organic_bosons.c:126-150), optional kick p += h_k F fused in the load;
//   filter:  z = a + i b of two links, Z' = g_n Z (real symmetric g), for the momentum refresh
//            g = sqrt(f_n) and the free-boson draw g = 1/sqrt(f_n) (organic_bosons.c:55-100);
//   energy:  sum_n |Z_n|^2 / (2 nt f_n) over the full spectrum = the two links' kinetic energies
//            (energy_fourier_momenta, organic_bosons.c:102-124).
// f_n = omega + (2 sin(pi n/nt))^2/omega.
#include "internal.h"

namespace ephmc {

constexpr int FFT_LW = 8;      // complex series per workgroup
constexpr int FFT_BT = 256;

enum FftMode { FFT_EVOLVE = 0, FFT_FILTER = 1, FFT_ENERGY = 2 };

struct FftArgs {
	const double *in0;     // evolve: x; filter/energy: input series
	const double *in1;     // evolve: p
	const double *kickF;   // evolve: force for the fused kick (or null)
	double kick;
	double *out0, *out1;   // evolve: x, p; filter: output series
	const double *twid;    // [nt] e^{-2 pi i j/nt} as (re, im)
	const double *f2;      // [nt] f_n
	double c, s;           // evolve: cos h / nt, sin h / nt
	int gain;              // filter: 0 -> sqrt(f)/nt, 1 -> 1/(sqrt(f) nt), 2 -> h/(f nt) added to out0
	double h;
	double *partial;       // energy: one value per workgroup
	unsigned loc, nt, logn;
	int pow2;
	int mode;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b){ return make_double2(a.x*b.x - a.y*b.y, a.x*b.y + a.y*b.x); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b){ return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b){ return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conjg(double2 a){ return make_double2(a.x, -a.y); }

// position of mode n in the transformed buffer
__device__ __forceinline__ unsigned mode_pos(unsigned n, const FftArgs &a)
{
	if(!a.pow2) return n;
	return a.logn ? (__brev(n) >> (32 - a.logn)) : 0u;
}

__device__ __forceinline__ void fa_fft_body(const FftArgs &a, unsigned bx)
{
	extern __shared__ __align__(16) unsigned char smem_raw[];
	const unsigned nt = a.nt;
	double2 *z = reinterpret_cast<double2 *>(smem_raw);            // [nt][LW]
	double2 *z2 = z + (size_t)nt * FFT_LW;                         // second buffer (direct sums)
	double2 *tw = z2 + (a.pow2 ? 0 : (size_t)nt * FFT_LW);         // [nt]
	__shared__ double red[FFT_BT];
	const int tid = threadIdx.x;
	const unsigned pairs = (a.mode == FFT_EVOLVE) ? 1 : 2;         // real series per complex slot
	const unsigned l0 = bx * FFT_LW * pairs;

	for(unsigned j = tid; j < nt; j += FFT_BT) tw[j] = make_double2(a.twid[2 * j], a.twid[2 * j + 1]);
	// ---- load
	for(unsigned e = tid; e < nt * FFT_LW; e += FFT_BT){
		const unsigned t = e / FFT_LW, l = e % FFT_LW;
		double re = 0, im = 0;
		if(a.mode == FFT_EVOLVE){
			const unsigned link = l0 + l;
			if(link < a.loc){
				const size_t q = (size_t)t * a.loc + link;
				re = a.in0[q];
				im = a.in1[q];
				if(a.kickF) im += a.kick * a.kickF[q];
			}
		}else{
			const unsigned la = l0 + l, lb = l0 + FFT_LW + l;
			if(la < a.loc) re = a.in0[(size_t)t * a.loc + la];
			if(lb < a.loc) im = a.in0[(size_t)t * a.loc + lb];
		}
		z[(size_t)t * FFT_LW + l] = make_double2(re, im);
	}
	__syncthreads();

	// ---- forward transform
	double2 *Z = z;
	if(a.pow2){
		for(unsigned half = nt >> 1; half >= 1; half >>= 1){
			const unsigned stride = nt / (2 * half);
			for(unsigned q = tid; q < (nt / 2) * FFT_LW; q += FFT_BT){
				const unsigned l = q % FFT_LW, b = q / FFT_LW;
				const unsigned grp = b / half, pos = b - grp * half;
				const unsigned i = grp * 2 * half + pos, j = i + half;
				const double2 u = z[(size_t)i * FFT_LW + l], v = z[(size_t)j * FFT_LW + l];
				z[(size_t)i * FFT_LW + l] = cadd(u, v);
				z[(size_t)j * FFT_LW + l] = cmul(csub(u, v), tw[pos * stride]);
			}
			__syncthreads();
		}
	}else{
		for(unsigned q = tid; q < nt * FFT_LW; q += FFT_BT){
			const unsigned l = q % FFT_LW, k = q / FFT_LW;
			double2 acc = make_double2(0, 0);
			for(unsigned t = 0; t < nt; t++) acc = cadd(acc, cmul(z[(size_t)t * FFT_LW + l], tw[((size_t)k * t) % nt]));
			z2[(size_t)k * FFT_LW + l] = acc;
		}
		__syncthreads();
		Z = z2;
	}

	// ---- per-mode operation
	if(a.mode == FFT_ENERGY){
		double e = 0;
		for(unsigned q = tid; q < nt * FFT_LW; q += FFT_BT){
			const unsigned l = q % FFT_LW, n = q / FFT_LW;
			const double2 v = Z[(size_t)mode_pos(n, a) * FFT_LW + l];
			e += (v.x * v.x + v.y * v.y) / (2.0 * nt * a.f2[n]);
		}
		red[tid] = e;
		__syncthreads();
		for(int off = FFT_BT / 2; off > 0; off >>= 1){
			if(tid < off) red[tid] += red[tid + off];
			__syncthreads();
		}
		if(tid == 0) a.partial[bx] = red[0];
		return;
	}
	if(a.mode == FFT_FILTER){
		for(unsigned q = tid; q < nt * FFT_LW; q += FFT_BT){
			const unsigned l = q % FFT_LW, n = q / FFT_LW;
			const double f = a.f2[n];
			const double g = a.gain == 0 ? sqrt(f) / nt : (a.gain == 1 ? 1.0 / (sqrt(f) * nt) : a.h / (f * nt));
			double2 &v = Z[(size_t)mode_pos(n, a) * FFT_LW + l];
			v = make_double2(g * v.x, g * v.y);
		}
	}else{
		for(unsigned q = tid; q < (nt / 2 + 1) * FFT_LW; q += FFT_BT){
			const unsigned l = q % FFT_LW, n = q / FFT_LW;
			const unsigned m = (nt - n) % nt;
			const unsigned pn = mode_pos(n, a), pm = mode_pos(m, a);
			const double2 zn = Z[(size_t)pn * FFT_LW + l], zm = conjg(Z[(size_t)pm * FFT_LW + l]);
			const double2 X = make_double2(0.5 * (zn.x + zm.x), 0.5 * (zn.y + zm.y));
			const double2 P = make_double2(0.5 * (zn.y - zm.y), -0.5 * (zn.x - zm.x));   // (zn - zm)/(2i)
			const double f = a.f2[n], sw = a.s * f, s1w = a.s / f;
			const double2 Xn = make_double2(a.c * X.x + s1w * P.x, a.c * X.y + s1w * P.y);
			const double2 Pn = make_double2(a.c * P.x - sw * X.x, a.c * P.y - sw * X.y);
			// Z'_n = X' + i P',  Z'_-n = conj X' + i conj P'
			Z[(size_t)pn * FFT_LW + l] = make_double2(Xn.x - Pn.y, Xn.y + Pn.x);
			if(m != n) Z[(size_t)pm * FFT_LW + l] = make_double2(Xn.x + Pn.y, -Xn.y + Pn.x);
		}
	}
	__syncthreads();

	// ---- inverse transform (unnormalised; the 1/nt is in the per-mode factors)
	if(a.pow2){
		for(unsigned half = 1; half <= nt / 2; half <<= 1){
			const unsigned stride = nt / (2 * half);
			for(unsigned q = tid; q < (nt / 2) * FFT_LW; q += FFT_BT){
				const unsigned l = q % FFT_LW, b = q / FFT_LW;
				const unsigned grp = b / half, pos = b - grp * half;
				const unsigned i = grp * 2 * half + pos, j = i + half;
				const double2 u = z[(size_t)i * FFT_LW + l];
				const double2 v = cmul(z[(size_t)j * FFT_LW + l], conjg(tw[pos * stride]));
				z[(size_t)i * FFT_LW + l] = cadd(u, v);
				z[(size_t)j * FFT_LW + l] = csub(u, v);
			}
			__syncthreads();
		}
	}else{
		for(unsigned q = tid; q < nt * FFT_LW; q += FFT_BT){
			const unsigned l = q % FFT_LW, t = q / FFT_LW;
			double2 acc = make_double2(0, 0);
			for(unsigned k = 0; k < nt; k++) acc = cadd(acc, cmul(z2[(size_t)k * FFT_LW + l], conjg(tw[((size_t)k * t) % nt])));
			z[(size_t)t * FFT_LW + l] = acc;
		}
		__syncthreads();
	}

	// ---- store
	for(unsigned e = tid; e < nt * FFT_LW; e += FFT_BT){
		const unsigned t = e / FFT_LW, l = e % FFT_LW;
		const double2 v = z[(size_t)t * FFT_LW + l];
		if(a.mode == FFT_EVOLVE){
			const unsigned link = l0 + l;
			if(link < a.loc){
				const size_t q = (size_t)t * a.loc + link;
				a.out0[q] = v.x;
				a.out1[q] = v.y;
			}
		}else{
			const unsigned la = l0 + l, lb = l0 + FFT_LW + l;
			if(a.gain == 2){
				if(la < a.loc) a.out0[(size_t)t * a.loc + la] += v.x;
				if(lb < a.loc) a.out0[(size_t)t * a.loc + lb] += v.y;
			}else{
				if(la < a.loc) a.out0[(size_t)t * a.loc + la] = v.x;
				if(lb < a.loc) a.out0[(size_t)t * a.loc + lb] = v.y;
			}
		}
	}
}

__global__ __launch_bounds__(FFT_BT) void fa_fft_kernel(FftArgs a)
{
	fa_fft_body(a, blockIdx.x);
}

// many chains in one launch: blockIdx.z = chain, with its own number of workgroups in gx
struct FftMultiArgs {
	FftArgs a;
	unsigned gx;
};

__global__ __launch_bounds__(FFT_BT) void fa_fft_multi_kernel(const FftMultiArgs *__restrict__ tab)
{
	const FftMultiArgs &m = tab[blockIdx.z];
	if(blockIdx.x >= m.gx) return;
	fa_fft_body(m.a, blockIdx.x);
}

static FftArgs fft_args(ephmc_ctx *c, int mode)
{
	FftArgs a{};
	a.twid = c->d_twiddle;
	a.f2 = c->d_f2mom;
	a.loc = c->loc;
	a.nt = c->nt;
	a.pow2 = (c->nt & (c->nt - 1)) == 0;
	unsigned lg = 0;
	while((1u << lg) < c->nt) lg++;
	a.logn = lg;
	a.mode = mode;
	return a;
}

// workgroups and LDS bytes of one transform launch
static int fft_shape(ephmc_ctx *c, const FftArgs &a, unsigned *blocks, size_t *lds)
{
	const unsigned per = (a.mode == FFT_EVOLVE) ? FFT_LW : 2 * FFT_LW;
	*blocks = (c->loc + per - 1) / per;
	*lds = (size_t)c->nt * FFT_LW * sizeof(double2) * (a.pow2 ? 1 : 2) + (size_t)c->nt * sizeof(double2);
	if(*lds > 160 * 1024){ set_error("fft: nt = %u needs %zu B of LDS", c->nt, *lds); return EPHMC_ERR_UNSUPPORTED; }
	return EPHMC_OK;
}

static int launch_fft(ephmc_ctx *c, const FftArgs &a)
{
	unsigned blocks;
	size_t lds;
	EPHMC_TRY(fft_shape(c, a, &blocks, &lds));
	(void)hipFuncSetAttribute((const void *)fa_fft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	fa_fft_kernel<<<blocks, FFT_BT, lds, c->stream>>>(a);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

static FftArgs evolve_args(ephmc_ctx *c, double *x, double *p, double h, const double *pdot, double kick)
{
	FftArgs a = fft_args(c, FFT_EVOLVE);
	a.f2 = c->d_f2;
	a.in0 = x; a.in1 = p; a.out0 = x; a.out1 = p;
	a.kickF = pdot; a.kick = kick;
	a.c = cos(h) / c->nt;
	a.s = sin(h) / c->nt;
	return a;
}

int fa_evolve(ephmc_ctx *c, double *x, double *p, double h, const double *pdot, double kick)
{
	return launch_fft(c, evolve_args(c, x, p, h, pdot, kick));
}

// inverse_sqrt = 0: momentum refresh with the kinetic frequencies (organic_bosons.c:76-100,
// omega -> sqrt(cfa_mass^2 + omega^2) under CLASSICAL_FOURIER_ACC); 1: free-boson draw with the
// bare frequencies (:55-74)
int fa_filter(ephmc_ctx *c, const double *in, double *out, int inverse_sqrt)
{
	FftArgs a = fft_args(c, FFT_FILTER);
	a.in0 = in; a.out0 = out;
	a.gain = inverse_sqrt ? 1 : 0;
	if(inverse_sqrt) a.f2 = c->d_f2;
	return launch_fft(c, a);
}

// x += c2r((h / f_n) r2c(p)) / nt with the kinetic frequencies: the drift of
// evolve_bosons_approx_step (organic_bosons.c:152-177)
int fa_drift(ephmc_ctx *c, const double *p, double *x, double h)
{
	FftArgs a = fft_args(c, FFT_FILTER);
	a.in0 = p; a.out0 = x;
	a.gain = 2;
	a.h = h;
	return launch_fft(c, a);
}

// boson_force (organic_bosons.c:179-194): pdot (+)= -omega x + (x_{t+1} - 2 x_t + x_{t-1}) / omega
__global__ void boson_force_kernel(const double *__restrict__ x, double *__restrict__ pdot, double omega, unsigned loc,
		size_t dim, int overwrite)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < dim; e += (size_t)gridDim.x * blockDim.x){
		const size_t nxt = (e + loc < dim) ? e + loc : e + loc - dim;
		const size_t prv = (e >= loc) ? e - loc : e + dim - loc;
		const double v = x[e];
		double f = overwrite ? 0.0 : pdot[e];
		f -= omega * v;
		f += (x[nxt] - 2 * v + x[prv]) / omega;
		pdot[e] = f;
	}
}

int boson_force(ephmc_ctx *c, const double *x, double *pdot, int overwrite)
{
	boson_force_kernel<<<1024, 256, 0, c->stream>>>(x, pdot, c->cfg.omega, c->loc, c->dim, overwrite);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

__global__ void axpy_kernel(double *__restrict__ y, double a, const double *__restrict__ x, size_t n)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) y[e] += a * x[e];
}

int axpy(ephmc_ctx *c, double *y, double a, const double *x, size_t n)
{
	axpy_kernel<<<1024, 256, 0, c->stream>>>(y, a, x, n);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// NO_FOURIER_ACC kinetic energy .5 * dim * average_sq(p, dim) (organic_hmc.c:57-59)
__global__ __launch_bounds__(256) void sumsq_kernel(const double *__restrict__ p, size_t n, double *__restrict__ part)
{
	__shared__ double red[256];
	double s = 0;
	for(size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) s += p[e] * p[e];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void plain_energy_final(const double *__restrict__ part, unsigned nb, double dim, double *__restrict__ out)
{
	__shared__ double red[256];
	double s = 0;
	for(unsigned i = threadIdx.x; i < nb; i += 256) s += part[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x == 0) *out = .5 * dim * (red[0] / dim);
}

int plain_energy(ephmc_ctx *c, const double *p, double *scal_out)
{
	sumsq_kernel<<<c->red_blocks, 256, 0, c->stream>>>(p, c->dim, c->d_red);
	EPHMC_LAUNCHED();
	plain_energy_final<<<1, 256, 0, c->stream>>>(c->d_red, c->red_blocks, (double)c->dim, scal_out);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

__global__ void sum_partials_kernel(const double *__restrict__ part, unsigned n, double *__restrict__ out)
{
	__shared__ double red[256];
	double s = 0;
	for(unsigned i = threadIdx.x; i < n; i += 256) s += part[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x == 0) *out = red[0];
}

int fa_energy(ephmc_ctx *c, const double *p, double *scal_out)
{
	FftArgs a = fft_args(c, FFT_ENERGY);
	a.in0 = p;
	a.partial = c->d_red;
	EPHMC_TRY(launch_fft(c, a));
	const unsigned blocks = (c->loc + 2 * FFT_LW - 1) / (2 * FFT_LW);
	sum_partials_kernel<<<1, 256, 0, c->stream>>>(c->d_red, blocks, scal_out);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// S_B(x) = (sum (x_{t+1}-x_t)^2 / omega + omega sum x^2) / 2   (hamilton, organic_hmc.c:67-84)
__global__ __launch_bounds__(256) void boson_action_kernel(const double *__restrict__ x, size_t dim, unsigned loc,
		double *__restrict__ part)
{
	__shared__ double rp[256], rk[256];
	double pot = 0, kin = 0;
	for(size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < dim; e += (size_t)gridDim.x * 256){
		const double v = x[e];
		const size_t nxt = (e + loc < dim) ? e + loc : e + loc - dim;
		const double d = x[nxt] - v;
		pot += v * v;
		kin += d * d;
	}
	rp[threadIdx.x] = pot;
	rk[threadIdx.x] = kin;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off){ rp[threadIdx.x] += rp[threadIdx.x + off]; rk[threadIdx.x] += rk[threadIdx.x + off]; }
		__syncthreads();
	}
	if(threadIdx.x == 0){ part[2 * blockIdx.x] = rp[0]; part[2 * blockIdx.x + 1] = rk[0]; }
}

__global__ void boson_action_final(const double *__restrict__ part, unsigned n, double omega, double *__restrict__ out)
{
	__shared__ double rp[256], rk[256];
	double pot = 0, kin = 0;
	for(unsigned i = threadIdx.x; i < n; i += 256){ pot += part[2 * i]; kin += part[2 * i + 1]; }
	rp[threadIdx.x] = pot;
	rk[threadIdx.x] = kin;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off){ rp[threadIdx.x] += rp[threadIdx.x + off]; rk[threadIdx.x] += rk[threadIdx.x + off]; }
		__syncthreads();
	}
	if(threadIdx.x == 0) *out = (rk[0] / omega + rp[0] * omega) / 2;
}

int boson_action(ephmc_ctx *c, const double *x, double *scal_out)
{
	const unsigned blocks = c->red_blocks;
	boson_action_kernel<<<blocks, 256, 0, c->stream>>>(x, c->dim, c->loc, c->d_red);
	EPHMC_LAUNCHED();
	boson_action_final<<<1, 256, 0, c->stream>>>(c->d_red, blocks, c->cfg.omega, scal_out);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// H = -flavors logdet + E_p + S_B into scal[out_slot]
__global__ void hamilton_kernel(double *scal, int logdet_slot, int ep_slot, int sb_slot, int out_slot, double flavors)
{
	scal[out_slot] = -scal[logdet_slot] * flavors + scal[ep_slot] + scal[sb_slot];
}

// energy_momenta (organic_hmc.c:56-65)
int kinetic_energy(ephmc_ctx *c, const double *p, double *scal_out)
{
	if(c->cfg.no_fourier_acc) return plain_energy(c, p, scal_out);
	return fa_energy(c, p, scal_out);
}

// H = -flavors logdet + E_p + S_B(x); p null: E_p = 0 (ZERO_FILLING, organic_hmc.c:252)
int hamilton_dev(ephmc_ctx *c, const double *x, const double *p, int logdet_slot, int out_slot)
{
	if(p) EPHMC_TRY(kinetic_energy(c, p, c->d_scal + SCAL_EP));
	else EPHMC_TRY(set_scalar(c, SCAL_EP, 0.0));
	EPHMC_TRY(boson_action(c, x, c->d_scal + SCAL_BOSON));
	hamilton_kernel<<<1, 1, 0, c->stream>>>(c->d_scal, logdet_slot, SCAL_EP, SCAL_BOSON, out_slot, (double)c->cfg.flavors);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// Metropolis (organic_hmc.c:233-243): accept iff exp(H_old - H_new) > U, NaN rejects.
__global__ void metropolis_kernel(double *scal, double uniform)
{
	const double hold = scal[SCAL_H_OLD], hnew = scal[SCAL_H_NEW];
	const double boltz = exp(hold - hnew);
	const int acc = (boltz > uniform) ? 1 : 0;
	scal[SCAL_ACC] = acc;
	scal[SCAL_BOLTZ] = boltz;
	scal[SCAL_DH] = hnew - hold;
	scal[SCAL_ENERGY] = acc ? hnew : hold;
	if(acc) scal[SCAL_LOGDET_CHAIN] = scal[SCAL_LOGDET];
}

// dst <- src when the proposal was accepted
__global__ void accept_copy_kernel(const double *__restrict__ scal, const double *__restrict__ src, double *__restrict__ dst, size_t n)
{
	if(scal[SCAL_ACC] == 0) return;
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) dst[e] = src[e];
}

int metropolis(ephmc_ctx *c, double uniform)
{
	metropolis_kernel<<<1, 1, 0, c->stream>>>(c->d_scal, uniform);
	EPHMC_LAUNCHED();
	accept_copy_kernel<<<1024, 256, 0, c->stream>>>(c->d_scal, c->d_xnew, c->d_x, c->dim);
	EPHMC_LAUNCHED();
	accept_copy_kernel<<<1024, 256, 0, c->stream>>>(c->d_scal, c->d_mhat, c->d_minv, c->mat);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// ---- many chains in one launch (batch.hip): the kernels above with the arguments of chain
// blockIdx.z read from a table.  The reductions keep each chain's own grid (gx), so that every
// partial sum covers the same elements as in the single-chain launch.
static int batch_fft(BatchOp &op, ephmc_ctx *c, const FftArgs &a)
{
	FftMultiArgs m{a, 0};
	size_t lds;
	EPHMC_TRY(fft_shape(c, a, &m.gx, &lds));
	(void)hipFuncSetAttribute((const void *)fa_fft_multi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
		(int)std::max(lds, op.lds));
	return op.add((const void *)fa_fft_multi_kernel, FFT_BT, lds, m.gx, 1, m);
}

int batch_fa_evolve(BatchOp &op, ephmc_ctx *c, double *x, double *p, double h, const double *pdot, double kick)
{
	return batch_fft(op, c, evolve_args(c, x, p, h, pdot, kick));
}

int batch_fa_filter(BatchOp &op, ephmc_ctx *c, const double *in, double *out)
{
	FftArgs a = fft_args(c, FFT_FILTER);
	a.in0 = in; a.out0 = out;
	a.gain = 0;
	return batch_fft(op, c, a);
}

struct SumArgs {
	const double *x;         // boson action: the field; sums: the partials
	size_t dim;
	unsigned loc, gx;
	double *part;            // boson action: partials
	double omega;
	double *out;
};

__global__ __launch_bounds__(256) void boson_action_multi_kernel(const SumArgs *__restrict__ tab)
{
	__shared__ double rp[256], rk[256];
	const SumArgs &a = tab[blockIdx.z];
	if(blockIdx.x >= a.gx) return;
	double pot = 0, kin = 0;
	for(size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < a.dim; e += (size_t)a.gx * 256){
		const double v = a.x[e];
		const size_t nxt = (e + a.loc < a.dim) ? e + a.loc : e + a.loc - a.dim;
		const double d = a.x[nxt] - v;
		pot += v * v;
		kin += d * d;
	}
	rp[threadIdx.x] = pot;
	rk[threadIdx.x] = kin;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off){ rp[threadIdx.x] += rp[threadIdx.x + off]; rk[threadIdx.x] += rk[threadIdx.x + off]; }
		__syncthreads();
	}
	if(threadIdx.x == 0){ a.part[2 * blockIdx.x] = rp[0]; a.part[2 * blockIdx.x + 1] = rk[0]; }
}

__global__ __launch_bounds__(256) void boson_action_final_multi_kernel(const SumArgs *__restrict__ tab)
{
	const SumArgs &a = tab[blockIdx.z];
	__shared__ double rp[256], rk[256];
	double pot = 0, kin = 0;
	for(unsigned i = threadIdx.x; i < a.gx; i += 256){ pot += a.x[2 * i]; kin += a.x[2 * i + 1]; }
	rp[threadIdx.x] = pot;
	rk[threadIdx.x] = kin;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off){ rp[threadIdx.x] += rp[threadIdx.x + off]; rk[threadIdx.x] += rk[threadIdx.x + off]; }
		__syncthreads();
	}
	if(threadIdx.x == 0) *a.out = (rk[0] / a.omega + rp[0] * a.omega) / 2;
}

__global__ __launch_bounds__(256) void sum_partials_multi_kernel(const SumArgs *__restrict__ tab)
{
	const SumArgs &a = tab[blockIdx.z];
	__shared__ double red[256];
	double s = 0;
	for(unsigned i = threadIdx.x; i < a.gx; i += 256) s += a.x[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x == 0) *a.out = red[0];
}

struct ScalArgs {
	double *scal;
	int logdet_slot, out_slot;
	double flavors;
	const double *uniform;
};

__global__ void hamilton_multi_kernel(const ScalArgs *__restrict__ tab)
{
	const ScalArgs &a = tab[blockIdx.z];
	a.scal[a.out_slot] = -a.scal[a.logdet_slot] * a.flavors + a.scal[SCAL_EP] + a.scal[SCAL_BOSON];
}

__global__ void metropolis_multi_kernel(const ScalArgs *__restrict__ tab)
{
	const ScalArgs &a = tab[blockIdx.z];
	double *scal = a.scal;
	const double hold = scal[SCAL_H_OLD], hnew = scal[SCAL_H_NEW];
	const double boltz = exp(hold - hnew);
	const int acc = (boltz > *a.uniform) ? 1 : 0;
	scal[SCAL_ACC] = acc;
	scal[SCAL_BOLTZ] = boltz;
	scal[SCAL_DH] = hnew - hold;
	scal[SCAL_ENERGY] = acc ? hnew : hold;
	if(acc) scal[SCAL_LOGDET_CHAIN] = scal[SCAL_LOGDET];
}

// hamilton_dev(c, x, p, logdet_slot, out_slot) with p != null in the Fourier-accelerated form:
// kinetic energy (transform, partial sum), boson action (partials, final), H
int batch_hamilton(BatchOp *ops, ephmc_ctx *c, const double *x, const double *p, int logdet_slot, int out_slot)
{
	if(c->cfg.no_fourier_acc){ set_error("batch: NO_FOURIER_ACC is not batched"); return EPHMC_ERR_UNSUPPORTED; }
	FftArgs a = fft_args(c, FFT_ENERGY);
	a.in0 = p;
	a.partial = c->d_red;
	EPHMC_TRY(batch_fft(ops[0], c, a));
	SumArgs s{};
	s.x = c->d_red;
	s.gx = (c->loc + 2 * FFT_LW - 1) / (2 * FFT_LW);
	s.out = c->d_scal + SCAL_EP;
	EPHMC_TRY(ops[1].add((const void *)sum_partials_multi_kernel, 256, 0, 1, 1, s));
	SumArgs b{};
	b.x = x;
	b.dim = c->dim;
	b.loc = c->loc;
	b.gx = c->red_blocks;
	b.part = c->d_red;
	EPHMC_TRY(ops[2].add((const void *)boson_action_multi_kernel, 256, 0, b.gx, 1, b));
	SumArgs f{};
	f.x = c->d_red;
	f.gx = c->red_blocks;
	f.omega = c->cfg.omega;
	f.out = c->d_scal + SCAL_BOSON;
	EPHMC_TRY(ops[3].add((const void *)boson_action_final_multi_kernel, 256, 0, 1, 1, f));
	const ScalArgs h{c->d_scal, logdet_slot, out_slot, (double)c->cfg.flavors, nullptr};
	return ops[4].add((const void *)hamilton_multi_kernel, 1, 0, 1, 1, h);
}

// metropolis(c, u) with u read from device memory (the accept copies are batch_copy records)
int batch_metropolis(BatchOp &op, ephmc_ctx *c, const double *uniform)
{
	const ScalArgs a{c->d_scal, 0, 0, 0.0, uniform};
	return op.add((const void *)metropolis_multi_kernel, 1, 0, 1, 1, a);
}

} // namespace ephmc
