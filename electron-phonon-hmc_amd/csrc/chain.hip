// chain.hip -- slice coefficients and the chained slice product on column blocks.
//
// A slice factor B_tau = p(A_tau) = prod_pairs (1 + a_lin A + a_quad A^2) with
// A_tau = mu + K(x_tau) the symmetric 7-point stencil of the triangular lattice
// (apply_mu_K_to_v, organic_dense_matrix.c:96-120; pair expansion :168-207).  The reference
// applies p(A) as stencils to every column of a dense N x N matrix (:209-216); this file does
// the same on the GPU: one workgroup owns C columns for the whole chain, keeps them in LDS
// (layout [site][column] so a neighbour read fetches C adjacent values) and applies all
// slices in sequence.  Each output site gathers its six bond neighbours:
//   w[i] = mu v[i] + sum_d coef[tau][d][i] v[nb(i,d)],  d = R, L, TR, BL, TL, BR,
// where coef[tau][d][i] = t_k - kappa_k x[tau][link(i,d)] is the coefficient of the bond
// between i and nb(i,d) (k = d/2).  This gather is exactly the reference's scatter over
// forward bonds, including lattices of extent 1 or 2 where bonds alias.
#include "internal.h"
#include "stencil.h"
#include <cstdlib>
#include <algorithm>
#include <cstring>

namespace ephmc {

__device__ __forceinline__ void coef_body(const double *__restrict__ x, const unsigned *__restrict__ link,
		double t0, double t1, double t2, double k0, double k1, double k2,
		unsigned ns, unsigned nt, double *__restrict__ c64, float *__restrict__ c32, unsigned bx, unsigned gx)
{
	const size_t n = (size_t)nt * 6 * ns;
	for(size_t e = bx * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gx * blockDim.x){
		const size_t tau = e / (6 * ns);
		const unsigned r = (unsigned)(e - tau * 6 * ns);
		const unsigned d = r / ns, s = r - d * ns;
		const unsigned k = d >> 1;
		const double tk = k == 0 ? t0 : (k == 1 ? t1 : t2);
		const double kk = k == 0 ? k0 : (k == 1 ? k1 : k2);
		const double f = tk - kk * x[tau * 3 * ns + link[d * ns + s]];
		if(c64) c64[e] = f;
		if(c32) c32[e] = (float)f;
	}
}

__global__ void coef_kernel(const double *__restrict__ x, const unsigned *__restrict__ link,
		double t0, double t1, double t2, double k0, double k1, double k2,
		unsigned ns, unsigned nt, double *__restrict__ c64, float *__restrict__ c32)
{
	coef_body(x, link, t0, t1, t2, k0, k1, k2, ns, nt, c64, c32, blockIdx.x, gridDim.x);
}

// many chains in one launch (blockIdx.z = chain; each chain's own grid in gx)
struct CoefArgs {
	const double *x;
	const unsigned *link;
	double t[3], k[3];
	unsigned ns, nt, gx;
	double *c64;
	float *c32;
};

__global__ void coef_multi_kernel(const CoefArgs *__restrict__ tab)
{
	const CoefArgs &a = tab[blockIdx.z];
	if(blockIdx.x >= a.gx) return;
	coef_body(a.x, a.link, a.t[0], a.t[1], a.t[2], a.k[0], a.k[1], a.k[2], a.ns, a.nt, a.c64, a.c32, blockIdx.x, a.gx);
}

int coefficients(ephmc_ctx *c, const double *x, bool want32)
{
	const size_t n = (size_t)c->nt * 6 * c->ns;
	const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
	coef_kernel<<<blocks, 256, 0, c->stream>>>(x, c->d_link, c->cfg.t[0], c->cfg.t[1], c->cfg.t[2],
		c->cfg.kappa[0], c->cfg.kappa[1], c->cfg.kappa[2], c->ns, c->nt, c->d_coef64, want32 ? c->d_coef32 : nullptr);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// ---------------------------------------------------------------------------------------------
template <typename R>
struct ChainArgs {
	const R *coef;          // [nt][6][ns]
	const int *nb;          // [6][ns]
	R mu;
	unsigned ns;
	const R *src;           // start columns (column-major), or null for the identity
	R *dst;                 // result columns
	R *ck;                  // checkpoints: state before every ck_every-th slice, or null
	unsigned ck_every;
	int tau0, dir, ntau;    // slices tau0, tau0+dir, ... (ntau of them)
	int np;
	R a_lin[MAX_PAIRS], a_quad[MAX_PAIRS];
	int cplx[MAX_PAIRS];
};

// out = mu*own + sum_d cf[d] * buf[nb[d]][c]
template <typename R, int C>
__device__ __forceinline__ void stencil_gather(const R *__restrict__ buf, const int (&nbr)[6], const R (&cf)[6],
		R mu, const R (&own)[C], R (&out)[C])
{
#pragma unroll
	for(int c = 0; c < C; c++) out[c] = mu * own[c];
	if constexpr(sizeof(R) == 4){
		// fp32: the six neighbour reads are issued together (fp64 keeps them interleaved with the
		// sums, which holds fewer registers); the sum is taken in the order d = 0..5 either way
		R nv[6][C];
#pragma unroll
		for(int d = 0; d < 6; d++)
#pragma unroll
			for(int c = 0; c < C; c++) nv[d][c] = buf[(size_t)nbr[d] * C + c];
#pragma unroll
		for(int d = 0; d < 6; d++)
#pragma unroll
			for(int c = 0; c < C; c++) out[c] = fmac(cf[d], nv[d][c], out[c]);
	}else{
#pragma unroll
		for(int d = 0; d < 6; d++){
			const R *q = buf + (size_t)nbr[d] * C;
#pragma unroll
			for(int c = 0; c < C; c++) out[c] = fmac(cf[d], q[c], out[c]);
		}
	}
}

// fp32 with one site per thread: every lane loads and gathers (a lane past the last site reads
// site ns-1 or neighbour 0, all in bounds, and stores nothing), so no lane needs zero values of
// its own, and the second LDS buffer sits at the constant offset C * BT, which a read takes as an
// immediate offset
template <typename R, int S>
constexpr bool chain_all_lanes(){ return sizeof(R) == 4 && S == 1; }

template <typename R, int C, int BT, int S>
constexpr size_t chain_lds_bytes(unsigned ns)
{
	return 2 * (chain_all_lanes<R, S>() ? (size_t)C * BT : (size_t)C * ns) * sizeof(R);
}

template <typename R, int C, int BT, int S>
__device__ __forceinline__ void chain_body(const ChainArgs<R> &a, unsigned bx)
{
	extern __shared__ __align__(16) unsigned char smem_raw[];
	constexpr bool ALL_LANES = chain_all_lanes<R, S>();
	const unsigned ns = a.ns;
	R *V = reinterpret_cast<R *>(smem_raw);
	R *W = V + (ALL_LANES ? (size_t)C * BT : (size_t)C * ns);
	const unsigned col0 = bx * C;
	const size_t mat = (size_t)ns * ns;

	unsigned site[S];
	bool act[S];
	int nbr[S][6];
	R v[S][C], w[S][C], cf[S][6];
#pragma unroll
	for(int j = 0; j < S; j++){
		site[j] = threadIdx.x + j * BT;
		act[j] = site[j] < ns;
#pragma unroll
		for(int d = 0; d < 6; d++) nbr[j][d] = act[j] ? a.nb[d * ns + site[j]] : 0;
#pragma unroll
		for(int c = 0; c < C; c++){
			const unsigned col = col0 + c;
			R val = 0;
			if(act[j] && col < ns) val = a.src ? a.src[(size_t)col * ns + site[j]] : (R)(site[j] == col ? 1 : 0);
			v[j][c] = val;
		}
	}

	// coefficients of the next slice are loaded one slice ahead, so their latency overlaps the
	// stencils of the current one
	R cfn[S][6];
	auto load_cf = [&](int i){
		const R *cft = a.coef + (size_t)(a.tau0 + a.dir * i) * 6 * ns;
#pragma unroll
		for(int j = 0; j < S; j++){
			const unsigned sl = ALL_LANES && !act[j] ? ns - 1 : site[j];
#pragma unroll
			for(int d = 0; d < 6; d++) cfn[j][d] = (ALL_LANES || act[j]) ? cft[d * ns + sl] : (R)0;
		}
	};
	if(a.ntau > 0) load_cf(0);
	for(int i = 0; i < a.ntau; i++){
		if(a.ck && (i % a.ck_every) == 0){
			R *ck = a.ck + (size_t)(i / a.ck_every) * mat;
#pragma unroll
			for(int j = 0; j < S; j++)
#pragma unroll
				for(int c = 0; c < C; c++)
					if(act[j] && col0 + c < ns) ck[(size_t)(col0 + c) * ns + site[j]] = v[j][c];
		}
#pragma unroll
		for(int j = 0; j < S; j++)
#pragma unroll
			for(int d = 0; d < 6; d++) cf[j][d] = cfn[j][d];
		if(i + 1 < a.ntau) load_cf(i + 1);

#pragma unroll
		for(int p = 0; p < MAX_PAIRS; p++){
			if(p >= a.np) break;
#pragma unroll
			for(int j = 0; j < S; j++)
				if(act[j])
#pragma unroll
					for(int c = 0; c < C; c++) V[(size_t)site[j] * C + c] = v[j][c];
			__syncthreads();
#pragma unroll
			for(int j = 0; j < S; j++)
				if(ALL_LANES || act[j]) stencil_gather<R, C>(V, nbr[j], cf[j], a.mu, v[j], w[j]);
			if(a.cplx[p]){
				// W is free: its last readers passed the barrier above.  The barrier below also
				// retires every read of V before the next pair overwrites it.
#pragma unroll
				for(int j = 0; j < S; j++)
					if(act[j])
#pragma unroll
						for(int c = 0; c < C; c++) W[(size_t)site[j] * C + c] = w[j][c];
				__syncthreads();
#pragma unroll
				for(int j = 0; j < S; j++){
					if(!ALL_LANES && !act[j]) continue;
					R z[C];
					stencil_gather<R, C>(W, nbr[j], cf[j], a.mu, w[j], z);
#pragma unroll
					for(int c = 0; c < C; c++) v[j][c] = v[j][c] + fmac(a.a_quad[p], z[c], a.a_lin[p] * w[j][c]);
				}
			}else{
				__syncthreads();
#pragma unroll
				for(int j = 0; j < S; j++)
#pragma unroll
					for(int c = 0; c < C; c++) v[j][c] = fmac(a.a_lin[p], w[j][c], v[j][c]);
			}
		}
	}
	if(a.dst){
#pragma unroll
		for(int j = 0; j < S; j++)
#pragma unroll
			for(int c = 0; c < C; c++)
				if(act[j] && col0 + c < ns) a.dst[(size_t)(col0 + c) * ns + site[j]] = v[j][c];
	}
}

template <typename R, int C, int BT, int S>
__global__ __launch_bounds__(BT) void chain_kernel(ChainArgs<R> a)
{
	chain_body<R, C, BT, S>(a, blockIdx.x);
}

// many chains in one launch: blockIdx.z = chain, blockIdx.x = its column block (blocks past the
// chain's own ceil(ns / C) return).  With one site per thread the workgroup may have fewer threads
// than BT (at least ns); the LDS layout stays that of BT.
template <typename R, int C, int BT, int S>
__global__ __launch_bounds__(BT) void chain_multi_kernel(const ChainArgs<R> *__restrict__ tab)
{
	const ChainArgs<R> &a = tab[blockIdx.z];
	if(blockIdx.x >= (a.ns + C - 1) / C) return;
	chain_body<R, C, BT, S>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------------------
// Strip form of the same chain (stencil.h): a thread owns M consecutive sites of a lattice row.
// The sum is taken in the same order (d = R, L, TR, BL, TL, BR) as stencil_gather, so both forms
// give identical bits.
template <typename R, int C, int M, int LB>
__global__ __launch_bounds__(LB) void chain_strip_kernel(ChainArgs<R> a)
{
	extern __shared__ __align__(16) unsigned char smem_raw[];
	R *V = reinterpret_cast<R *>(smem_raw);
	R *W = V + (size_t)C * a.ns;
	const unsigned ns = a.ns;
	const unsigned col0 = blockIdx.x * C;
	const size_t mat = (size_t)ns * ns;
	const unsigned s0 = threadIdx.x * M;      // blockDim.x * M == ns
	StripNb<M> sn;
	strip_neighbours<M>(a.nb, ns, s0, sn);
	R v[M][C], w[M][C], cf[M][6];
#pragma unroll
	for(int k = 0; k < M; k++)
#pragma unroll
		for(int c = 0; c < C; c++){
			const unsigned col = col0 + c, site = s0 + k;
			R val = 0;
			if(col < ns) val = a.src ? a.src[(size_t)col * ns + site] : (R)(site == col ? 1 : 0);
			v[k][c] = val;
		}
	for(int i = 0; i < a.ntau; i++){
		const int tau = a.tau0 + a.dir * i;
		if(a.ck && (i % a.ck_every) == 0){
			R *ck = a.ck + (size_t)(i / a.ck_every) * mat;
#pragma unroll
			for(int k = 0; k < M; k++)
#pragma unroll
				for(int c = 0; c < C; c++)
					if(col0 + c < ns) ck[(size_t)(col0 + c) * ns + s0 + k] = v[k][c];
		}
		const R *cft = a.coef + (size_t)tau * 6 * ns;
#pragma unroll
		for(int k = 0; k < M; k++)
#pragma unroll
			for(int d = 0; d < 6; d++) cf[k][d] = cft[d * ns + s0 + k];
		for(int p = 0; p < a.np; p++){
#pragma unroll
			for(int k = 0; k < M; k++)
#pragma unroll
				for(int c = 0; c < C; c++) V[(size_t)(s0 + k) * C + c] = v[k][c];
			__syncthreads();
			strip_gather<R, C, M>(V, sn, cf, a.mu, v, w);
			if(a.cplx[p]){
#pragma unroll
				for(int k = 0; k < M; k++)
#pragma unroll
					for(int c = 0; c < C; c++) W[(size_t)(s0 + k) * C + c] = w[k][c];
				__syncthreads();
				R z[M][C];
				strip_gather<R, C, M>(W, sn, cf, a.mu, w, z);
#pragma unroll
				for(int k = 0; k < M; k++)
#pragma unroll
					for(int c = 0; c < C; c++) v[k][c] = v[k][c] + fmac(a.a_quad[p], z[k][c], a.a_lin[p] * w[k][c]);
			}else{
				__syncthreads();
#pragma unroll
				for(int k = 0; k < M; k++)
#pragma unroll
					for(int c = 0; c < C; c++) v[k][c] = fmac(a.a_lin[p], w[k][c], v[k][c]);
			}
		}
	}
	if(a.dst){
#pragma unroll
		for(int k = 0; k < M; k++)
#pragma unroll
			for(int c = 0; c < C; c++)
				if(col0 + c < ns) a.dst[(size_t)(col0 + c) * ns + s0 + k] = v[k][c];
	}
}

// Strip width for this lattice and arithmetic (0: site-per-thread form): the width chosen by
// ephmc_autotune, else EPHMC_CHAIN_STRIP=0/2/4, else 2 in fp64 and 0 in fp32.  The strip form
// cuts the LDS reads by a third (M = 2) at a larger register footprint; at 32x32 it is the
// faster fp64 chain (0.53 against 0.76 ms per 128 slices) and the slower fp32 one (0.31 against
// 0.28 ms).
unsigned strip_width(const ephmc_ctx *c, int prec)
{
	const int slot = prec == EPHMC_FP64 ? 0 : 1;
	unsigned m = prec == EPHMC_FP64 ? 2 : 0;
	if(c->chain_strip_set[slot] >= 0) m = (unsigned)c->chain_strip_set[slot];
	else if(const char *e = getenv("EPHMC_CHAIN_STRIP")) m = (unsigned)atoi(e);
	if(m != 2 && m != 4) return 0;
	if(c->l1 % m != 0) m = (m == 4 && c->l1 % 2 == 0) ? 2 : 0;
	if(m == 0 || c->ns / m > 1024) return 0;
	return m;
}

template <typename R, int C, int M>
static int launch_strip(ephmc_ctx *c, const ChainArgs<R> &a)
{
	const unsigned ns = a.ns;
	const size_t lds = 2 * (size_t)C * ns * sizeof(R);
	if(lds > 160 * 1024){ set_error("chain: %zu B of LDS exceeds 160 KiB", lds); return EPHMC_ERR_ARG; }
	const unsigned bt = ns / M, blocks = (ns + C - 1) / C;
#define EPHMC_STRIP_LAUNCH(LB)                                                                          \
	do {                                                                                                \
		(void)hipFuncSetAttribute((const void *)chain_strip_kernel<R, C, M, LB>,                        \
			hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                                       \
		chain_strip_kernel<R, C, M, LB><<<blocks, bt, lds, c->stream>>>(a);                              \
	} while(0)
	if(bt <= 256) EPHMC_STRIP_LAUNCH(256);
	else if(bt <= 512) EPHMC_STRIP_LAUNCH(512);
	else EPHMC_STRIP_LAUNCH(1024);
#undef EPHMC_STRIP_LAUNCH
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// Sites per thread requested by a tuning variable (1 when unset): 2 runs lattices of 513..1024
// sites on 512 threads.
unsigned sites_per_thread(const char *var)
{
	const char *e = getenv(var);
	return (e && atoi(e) == 2) ? 2u : 1u;
}

// Launch shape: threads per workgroup and sites per thread from the lattice size; columns
// per workgroup from the tuning table (ctx->cols) capped by LDS.
template <typename R, int C>
static int launch_chain_c(ephmc_ctx *c, const ChainArgs<R> &a)
{
	const unsigned ns = a.ns;
	const unsigned blocks = (ns + C - 1) / C;
#define EPHMC_CHAIN_LAUNCH(BT, S)                                                                  \
	do {                                                                                           \
		const size_t lds = chain_lds_bytes<R, C, BT, S>(ns);                                       \
		if(lds > 160 * 1024){ set_error("chain: %zu B of LDS exceeds 160 KiB", lds); return EPHMC_ERR_ARG; } \
		(void)hipFuncSetAttribute((const void *)chain_kernel<R, C, BT, S>,                               \
			hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                                 \
		chain_kernel<R, C, BT, S><<<blocks, BT, lds, c->stream>>>(a);                             \
	} while(0)
	if(ns <= 256) EPHMC_CHAIN_LAUNCH(256, 1);
	else if(ns <= 512) EPHMC_CHAIN_LAUNCH(512, 1);
	else if(ns <= 1024 && sites_per_thread("EPHMC_CHAIN_SITES") == 2) EPHMC_CHAIN_LAUNCH(512, 2);
	else if(ns <= 1024) EPHMC_CHAIN_LAUNCH(1024, 1);
	else if(ns <= 2048) EPHMC_CHAIN_LAUNCH(1024, 2);
	else if(ns <= 4096) EPHMC_CHAIN_LAUNCH(1024, 4);
	else { set_error("chain: lattice of %u sites exceeds 4096", ns); return EPHMC_ERR_UNSUPPORTED; }
#undef EPHMC_CHAIN_LAUNCH
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

template <typename R>
static int launch_chain(ephmc_ctx *c, const ChainArgs<R> &a, unsigned cols)
{
	// LDS budget: 2 * C * ns * sizeof(R) <= 160 KiB
	while(cols > 1 && 2 * (size_t)cols * a.ns * sizeof(R) > 160 * 1024) cols >>= 1;
	const unsigned m = strip_width(c, sizeof(R) == 8 ? EPHMC_FP64 : EPHMC_MIXED);
	if(m == 4){
		switch(cols){
			case 1: return launch_strip<R, 1, 4>(c, a);
			case 2: return launch_strip<R, 2, 4>(c, a);
			case 4: return launch_strip<R, 4, 4>(c, a);
			default: return launch_strip<R, 8, 4>(c, a);
		}
	}
	if(m == 2){
		switch(cols){
			case 1: return launch_strip<R, 1, 2>(c, a);
			case 2: return launch_strip<R, 2, 2>(c, a);
			case 4: return launch_strip<R, 4, 2>(c, a);
			default: return launch_strip<R, 8, 2>(c, a);
		}
	}
	switch(cols){
		case 1: return launch_chain_c<R, 1>(c, a);
		case 2: return launch_chain_c<R, 2>(c, a);
		case 4: return launch_chain_c<R, 4>(c, a);
		default: return launch_chain_c<R, 8>(c, a);
	}
}

template <typename R>
static ChainArgs<R> make_args(ephmc_ctx *c, const R *coef)
{
	ChainArgs<R> a{};
	a.coef = coef;
	a.nb = c->d_nb;
	a.mu = (R)c->cfg.mu;
	a.ns = c->ns;
	a.np = c->poly.np;
	for(int p = 0; p < c->poly.np; p++){
		a.a_lin[p] = (R)c->poly.a_lin[p];
		a.a_quad[p] = (R)c->poly.a_quad[p];
		a.cplx[p] = c->poly.cplx[p];
	}
	return a;
}

unsigned chain_cols(const ephmc_ctx *c, int prec)
{
	// columns per workgroup of the forward chain: the widest that gives at least 512 workgroups
	// (two per CU); EPHMC_CHAIN_COLS overrides it for tuning sweeps
	const unsigned ns = c->ns;
	const int slot = prec == EPHMC_FP64 ? 0 : 1;
	if(c->chain_cols_set[slot]) return c->chain_cols_set[slot];
	if(const char *e = getenv("EPHMC_CHAIN_COLS")){
		const unsigned v = (unsigned)atoi(e);
		if(v == 1 || v == 2 || v == 4 || v == 8) return v;
	}
	unsigned cols = 8;
	while(cols > 1 && (ns + cols - 1) / cols < 512) cols >>= 1;
	return cols;
}

// Chain of slices over all columns: dst = B_{tau_last} ... B_{tau_first} src, with the
// slices tau0, tau0+dir, ... (nt of them); src null = identity.  Stores the state before
// every ck_every-th slice into ck.  prec selects fp64 or fp32 storage and arithmetic.
template <typename R>
static ChainArgs<R> forward_args(ephmc_ctx *c, const R *coef, const void *src, void *dst, void *ck, unsigned ck_every,
		int descending, bool from_identity)
{
	ChainArgs<R> a = make_args<R>(c, coef);
	a.src = from_identity ? nullptr : (const R *)src;
	a.dst = (R *)dst;
	a.ck = (R *)ck;
	a.ck_every = ck_every ? ck_every : 1;
	a.tau0 = descending ? (int)c->nt - 1 : 0;
	a.dir = descending ? -1 : 1;
	a.ntau = (int)c->nt;
	return a;
}

int chain_forward(ephmc_ctx *c, int prec, const void *src, void *dst, void *ck, unsigned ck_every,
		int descending, bool from_identity)
{
	if(prec == EPHMC_FP64)
		return launch_chain<double>(c, forward_args<double>(c, c->d_coef64, src, dst, ck, ck_every, descending, from_identity),
			chain_cols(c, prec));
	return launch_chain<float>(c, forward_args<float>(c, c->d_coef32, src, dst, ck, ck_every, descending, from_identity),
		chain_cols(c, prec));
}

// dst = B_{tau_last} ... B_{tau0} src over the slices tau0, tau0+dir, ... (ntau of them) in the
// precision prec; src is fp64 (converted for fp32) or null for the identity.
int chain_range(ephmc_ctx *c, int prec, const double *src, void *dst, int tau0, int dir, int ntau)
{
	if(prec == EPHMC_FP64){
		ChainArgs<double> a = make_args<double>(c, c->d_coef64);
		a.src = src;
		a.dst = (double *)dst;
		a.ck = nullptr;
		a.ck_every = 1;
		a.tau0 = tau0; a.dir = dir; a.ntau = ntau;
		return launch_chain<double>(c, a, chain_cols(c, prec));
	}
	float *s32 = nullptr;
	if(src){
		s32 = (float *)c->d_seg;        // scratch: the segment buffer holds at least N^2 values
		EPHMC_TRY(convert_f64_to_f32(c, src, s32, c->mat));
	}
	ChainArgs<float> a = make_args<float>(c, c->d_coef32);
	a.src = s32;
	a.dst = (float *)dst;
	a.ck = nullptr;
	a.ck_every = 1;
	a.tau0 = tau0; a.dir = dir; a.ntau = ntau;
	return launch_chain<float>(c, a, chain_cols(c, prec));
}

// fp64 coefficients of the single slice whose 3N couplings are x_tau, into d_coef64 slot 0
int slice_coefficients(ephmc_ctx *c, const double *x_tau)
{
	const unsigned ns = c->ns;
	coef_kernel<<<(6 * ns + 255) / 256, 256, 0, c->stream>>>(x_tau, c->d_link, c->cfg.t[0], c->cfg.t[1], c->cfg.t[2],
		c->cfg.kappa[0], c->cfg.kappa[1], c->cfg.kappa[2], ns, 1, c->d_coef64, nullptr);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// apply_exp_mu_K_to_M with partial pair products (revert/stop, organic_dense_matrix.c:186-194)
// for a single slice whose 3N couplings are x_tau.  Uses d_coef64 slot 0 as scratch.
int apply_exp_partial(ephmc_ctx *c, const double *x_tau, double *M, int revert, int stop)
{
	EPHMC_TRY(slice_coefficients(c, x_tau));
	ChainArgs<double> a = make_args<double>(c, c->d_coef64);
	a.np = 0;
	auto add = [&](int p){
		a.a_lin[a.np] = c->poly.a_lin[p];
		a.a_quad[a.np] = c->poly.a_quad[p];
		a.cplx[a.np] = c->poly.cplx[p];
		a.np++;
	};
	if(revert){ for(int p = c->poly.np - 1; p > stop; p--) add(p); }
	else{ for(int p = 0; p < stop; p++) add(p); }
	if(a.np == 0) return EPHMC_OK;
	a.src = M;
	a.dst = M;       // each workgroup reads its columns before writing them
	a.ck = nullptr;
	a.ck_every = 1;
	a.tau0 = 0; a.dir = 1; a.ntau = 1;
	return launch_chain<double>(c, a, chain_cols(c, EPHMC_FP64));
}

// ---------------------------------------------------------------------------------------------
// M = T^T (+ identity): 32x32 tiles through LDS; T in R, M in fp64.
template <typename R>
__device__ __forceinline__ void transpose_body(const R *__restrict__ T, double *__restrict__ M, unsigned n, int add_identity,
		unsigned bxi, unsigned byi)
{
	__shared__ double tile[32][33];
	const unsigned bx = bxi * 32, by = byi * 32;
	const unsigned tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for(unsigned r = ty; r < 32; r += 8){
		const unsigned col = by + r, row = bx + tx;          // T[col][row] (column col of T)
		if(col < n && row < n) tile[r][tx] = (double)T[(size_t)col * n + row];
	}
	__syncthreads();
	for(unsigned r = ty; r < 32; r += 8){
		const unsigned col = bx + r, row = by + tx;          // M column col = row col of T
		if(col < n && row < n) M[(size_t)col * n + row] = tile[tx][r] + ((add_identity && col == row) ? 1.0 : 0.0);
	}
}

template <typename R>
__global__ __launch_bounds__(256) void transpose_kernel(const R *__restrict__ T, double *__restrict__ M,
		unsigned n, int add_identity)
{
	transpose_body<R>(T, M, n, add_identity, blockIdx.x, blockIdx.y);
}

struct TransArgs {
	const void *T;
	double *M;
	unsigned n;
	int add_identity;
};

template <typename R>
__global__ __launch_bounds__(256) void transpose_multi_kernel(const TransArgs *__restrict__ tab)
{
	const TransArgs &a = tab[blockIdx.z];
	const unsigned g = (a.n + 31) / 32;
	if(blockIdx.x >= g || blockIdx.y >= g) return;
	transpose_body<R>((const R *)a.T, a.M, a.n, a.add_identity, blockIdx.x, blockIdx.y);
}

int transpose_plus_identity(ephmc_ctx *c, int prec, const void *T, double *M, int add_identity)
{
	const unsigned n = c->ns;
	dim3 grid((n + 31) / 32, (n + 31) / 32);
	if(prec == EPHMC_FP64) transpose_kernel<double><<<grid, 256, 0, c->stream>>>((const double *)T, M, n, add_identity);
	else transpose_kernel<float><<<grid, 256, 0, c->stream>>>((const float *)T, M, n, add_identity);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

__global__ void add_identity_kernel(double *__restrict__ M, unsigned n)
{
	for(unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) M[(size_t)i * (n + 1)] += 1.0;
}

struct AddIdArgs {
	double *M;
	unsigned n;
};

__global__ void add_identity_multi_kernel(const AddIdArgs *__restrict__ tab)
{
	const AddIdArgs &a = tab[blockIdx.z];
	for(unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) a.M[(size_t)i * (a.n + 1)] += 1.0;
}

int add_identity(ephmc_ctx *c, double *M)
{
	add_identity_kernel<<<(c->ns + 255) / 256, 256, 0, c->stream>>>(M, c->ns);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// SINGLE_PARTICLE log det = log tr(B_0 ... B_{nt-1}) (organic_dense_matrix.c:550-551); the force
// factor exp(-logdet) (organic_hmc.c:121) goes to fs_slot when >= 0.
template <typename R>
__global__ __launch_bounds__(256) void log_trace_kernel(const R *__restrict__ M, unsigned n, double *__restrict__ scal,
		int ld_slot, int fs_slot)
{
	__shared__ double red[256];
	double s = 0;
	for(unsigned i = threadIdx.x; i < n; i += 256) s += (double)M[(size_t)i * (n + 1)];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x == 0){
		const double ld = log(red[0]);
		scal[ld_slot] = ld;
		if(fs_slot >= 0) scal[fs_slot] = exp(-ld);
	}
}

int log_trace(ephmc_ctx *c, int prec, const void *M, int ld_slot, int fs_slot)
{
	if(prec == EPHMC_FP64) log_trace_kernel<double><<<1, 256, 0, c->stream>>>((const double *)M, c->ns, c->d_scal, ld_slot, fs_slot);
	else log_trace_kernel<float><<<1, 256, 0, c->stream>>>((const float *)M, c->ns, c->d_scal, ld_slot, fs_slot);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

template <typename R>
__global__ void set_identity_kernel(R *__restrict__ M, unsigned n)
{
	const size_t nn = (size_t)n * n;
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x)
		M[e] = (e % (n + 1) == 0) ? (R)1 : (R)0;
}

int set_identity(ephmc_ctx *c, int prec, void *M)
{
	if(prec == EPHMC_FP64) set_identity_kernel<double><<<1024, 256, 0, c->stream>>>((double *)M, c->ns);
	else set_identity_kernel<float><<<1024, 256, 0, c->stream>>>((float *)M, c->ns);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

__global__ void set_scalar_kernel(double *scal, int slot, double v){ scal[slot] = v; }

int set_scalar(ephmc_ctx *c, int slot, double v)
{
	set_scalar_kernel<<<1, 1, 0, c->stream>>>(c->d_scal, slot, v);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

__global__ void f64_to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, size_t n)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
		b[e] = (float)a[e];
}

int convert_f64_to_f32(ephmc_ctx *c, const double *a, float *b, size_t n)
{
	const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
	f64_to_f32_kernel<<<blocks, 256, 0, c->stream>>>(a, b, n);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// ---- many chains in one launch (batch.hip).  The chain runs in the site form with 2 columns per
// workgroup on `block` threads (ns <= block <= threads, the LDS layout's workgroup size); every
// column is computed with the arithmetic of every other form and width, so a chain gives the same
// bits in a batch as alone.
int batch_coef(BatchOp &op, ephmc_ctx *c, const double *x, bool want32)
{
	CoefArgs a{};
	a.x = x;
	a.link = c->d_link;
	for(int k = 0; k < 3; k++){ a.t[k] = c->cfg.t[k]; a.k[k] = c->cfg.kappa[k]; }
	a.ns = c->ns;
	a.nt = c->nt;
	const size_t n = (size_t)c->nt * 6 * c->ns;
	a.gx = (unsigned)std::min<size_t>((n + 255) / 256, 4096);     // the grid of coefficients()
	a.c64 = c->d_coef64;
	a.c32 = want32 ? c->d_coef32 : nullptr;
	return op.add((const void *)coef_multi_kernel, 256, 0, a.gx, 1, a);
}

template <typename R, int BT, int C>
static int batch_chain_c(BatchOp &op, unsigned block, const ChainArgs<R> &a)
{
	const size_t lds = chain_lds_bytes<R, C, BT, 1>(a.ns);
	if(lds > 160 * 1024){ set_error("batch chain: %zu B of LDS exceeds 160 KiB", lds); return EPHMC_ERR_ARG; }
	const void *fn = (const void *)chain_multi_kernel<R, C, BT, 1>;
	(void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max(lds, op.lds));
	return op.add(fn, block, lds, (a.ns + C - 1) / C, 1, a);
}

// columns per workgroup: 2, or EPHMC_BATCH_CHAIN_COLS = 1 / 4 for tuning sweeps (any width gives the
// same bits)
template <typename R, int BT>
static int batch_chain_t(BatchOp &op, unsigned block, const ChainArgs<R> &a)
{
	const char *e = getenv("EPHMC_BATCH_CHAIN_COLS");
	const int cols = e ? atoi(e) : 2;
	if(cols == 1) return batch_chain_c<R, BT, 1>(op, block, a);
	if(cols == 4) return batch_chain_c<R, BT, 4>(op, block, a);
	return batch_chain_c<R, BT, 2>(op, block, a);
}

template <typename R>
static int batch_chain_r(BatchOp &op, unsigned threads, unsigned block, const ChainArgs<R> &a)
{
	if(a.ns > block || block > threads || block % 64){
		set_error("batch chain: %u sites on %u of %u threads", a.ns, block, threads);
		return EPHMC_ERR_ARG;
	}
	switch(threads){
		case 256: return batch_chain_t<R, 256>(op, block, a);
		case 512: return batch_chain_t<R, 512>(op, block, a);
		case 1024: return batch_chain_t<R, 1024>(op, block, a);
		default: set_error("batch chain: %u threads per workgroup", threads); return EPHMC_ERR_ARG;
	}
}

int batch_chain(BatchOp &op, ephmc_ctx *c, int prec, unsigned threads, unsigned block, void *dst, void *ck,
		unsigned ck_every, int descending)
{
	if(prec == EPHMC_FP64)
		return batch_chain_r<double>(op, threads, block,
			forward_args<double>(c, c->d_coef64, nullptr, dst, ck, ck_every, descending, true));
	return batch_chain_r<float>(op, threads, block, forward_args<float>(c, c->d_coef32, nullptr, dst, ck, ck_every, descending, true));
}

int batch_transpose(BatchOp &op, ephmc_ctx *c, int prec, const void *T, double *M)
{
	const TransArgs a{T, M, c->ns, 1};
	const unsigned g = (c->ns + 31) / 32;
	const void *fn = prec == EPHMC_FP64 ? (const void *)transpose_multi_kernel<double> : (const void *)transpose_multi_kernel<float>;
	return op.add(fn, 256, 0, g, g, a);
}

int batch_add_identity(BatchOp &op, ephmc_ctx *c, double *M)
{
	const AddIdArgs a{M, c->ns};
	return op.add((const void *)add_identity_multi_kernel, 256, 0, (c->ns + 255) / 256, 1, a);
}

struct CopyArgs {
	const double *src;
	double *dst;
	size_t n;
	const double *gate;
};

// elementwise, so any grid gives the same result
__global__ __launch_bounds__(256) void copy_multi_kernel(const CopyArgs *__restrict__ tab)
{
	const CopyArgs &a = tab[blockIdx.z];
	if(a.gate && a.gate[SCAL_ACC] == 0) return;
	for(size_t e = blockIdx.x * (size_t)256 + threadIdx.x; e < a.n; e += (size_t)gridDim.x * 256) a.dst[e] = a.src[e];
}

int batch_copy(BatchOp &op, const double *src, double *dst, size_t n, const double *gate)
{
	const CopyArgs a{src, dst, n, gate};
	return op.add((const void *)copy_multi_kernel, 256, 0, (unsigned)std::min<size_t>((n + 255) / 256, 1024), 1, a);
}

} // namespace ephmc
