// correlator.hip -- current-current correlator of a measured configuration, replacing
// cc_correlator (organic_dense_matrix.c:451-524) on the GPU (FP64).
//
// Transfer matrices (same storage convention as the reference, column-major chunks):
//   R_t0[tau] = B_tau ... B_{nt-1} M_hat^{-1}          (slots 0..nt-1, built slice by slice
//               from nt-1 down to 0; slot tau keeps the state after slice tau)
//   R_0t[tau] = B_{tau-1} ... B_0 (M_hat^{-1})^T        (used at tau before slice tau is applied)
// and for every tau and pair of bond directions (a, b)
//   g_ab(tau) = 1/N sum_{i,j} chi_ia chi_jb [R_t0[j*N+i] R_0t[jb*N+ia] + R_t0[jb*N+ia] R_0t[j*N+i]
//                                        - R_t0[jb*N+i] R_0t[j*N+ia] - R_t0[j*N+ia] R_0t[jb*N+i]]
// with ia = nb(i, 2a), jb = nb(j, 2b), chi_ia = t_a - kappa_a x[0][i][a], chi_jb from slice tau.
// One launch covers every tau: a workgroup takes a chunk of j for one tau and, for each j, stages
// the rows j and jb (b = 0..2) of both propagators in LDS, from where its threads gather the
// entries at i and ia; block sums per (tau, j-chunk), then a fixed-order pass per tau.
#include "internal.h"
#include "alloc_group.h"
#include <mutex>

namespace ephmc {

constexpr int CC_BT = 1024;
constexpr unsigned CC_J = 8;   // values of j per workgroup

// block sums of one correlator call: [nt][ceil(N/CC_J)][9]
static size_t cc_partials(unsigned ns, unsigned nt)
{
	return (size_t)nt * ((ns + CC_J - 1) / CC_J) * 9;
}

// R_t0 and R_0t (store values each) and the result block (res values), all or none
static int cc_alloc(ephmc_ctx *c, size_t store, size_t res)
{
	if(c->d_Rt0) (void)hipFree(c->d_Rt0);
	if(c->d_R0t) (void)hipFree(c->d_R0t);
	if(c->d_ccg) (void)hipFree(c->d_ccg);
	void **ptrs[3] = {(void **)&c->d_Rt0, (void **)&c->d_R0t, (void **)&c->d_ccg};
	const size_t bytes[3] = {store * sizeof(double), store * sizeof(double), res * sizeof(double)};
	const bool ok = alloc_all_or_none(ptrs, bytes, 3,
		[](void **p, size_t b){ return hipMalloc(p, b) == hipSuccess; },
		[](void *p){ (void)hipFree(p); });
	if(!ok){
		(void)hipGetLastError();
		set_error("correlator buffers (%zu MB) could not be allocated", 2 * store * 8 >> 20);
		return EPHMC_ERR_ALLOC;
	}
	return EPHMC_OK;
}

// grid (j-chunks of CC_J, tau).  rows (dynamic LDS, 8 N doubles): R_t0[j], R_0t[j], R_t0[jb]
// (b = 0..2), R_0t[jb] (b = 0..2); thread t takes the sites i = t, t + CC_BT, ...
__global__ __launch_bounds__(CC_BT) void cc_contract_kernel(const double *__restrict__ Tfin, const double *__restrict__ Rt0,
		const double *__restrict__ R0t, const double *__restrict__ x, const int *__restrict__ nb, double t0, double t1, double t2,
		double k0, double k1, double k2, unsigned ns, unsigned nt, double *__restrict__ partial)
{
	extern __shared__ double rows[];
	__shared__ double red[9][CC_BT / 64];
	const double tk[3] = {t0, t1, t2}, kk[3] = {k0, k1, k2};
	const unsigned tau = blockIdx.y;
	const size_t mat = (size_t)ns * ns;
	const double *rt = tau == 0 ? Tfin : Rt0 + (size_t)(nt - tau) * mat;
	const double *r0 = R0t + (size_t)tau * mat;
	const double *xt = x + (size_t)tau * 3 * ns;
	double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
	const unsigned j0 = blockIdx.x * CC_J, j1 = min(j0 + CC_J, ns);
	for(unsigned j = j0; j < j1; j++){
		unsigned jb[3];
		double chijb[3];
#pragma unroll
		for(int b = 0; b < 3; b++){
			jb[b] = (unsigned)nb[(2 * b) * ns + j];
			chijb[b] = tk[b] - kk[b] * xt[(size_t)j * 3 + b];
		}
		__syncthreads();                 // the previous j's rows are no longer read
		for(unsigned i = threadIdx.x; i < ns; i += CC_BT){
			rows[i] = rt[(size_t)j * ns + i];
			rows[ns + i] = r0[(size_t)j * ns + i];
#pragma unroll
			for(int b = 0; b < 3; b++){
				rows[(2 + b) * ns + i] = rt[(size_t)jb[b] * ns + i];
				rows[(5 + b) * ns + i] = r0[(size_t)jb[b] * ns + i];
			}
		}
		__syncthreads();
		for(unsigned i = threadIdx.x; i < ns; i += CC_BT){
			const double rt_ji = rows[i], r0_ji = rows[ns + i];
			unsigned ia[3];
			double chia[3], rt_jia[3], r0_jia[3];
#pragma unroll
			for(int a = 0; a < 3; a++){
				ia[a] = (unsigned)nb[(2 * a) * ns + i];
				chia[a] = tk[a] - kk[a] * x[(size_t)i * 3 + a];
				rt_jia[a] = rows[ia[a]];
				r0_jia[a] = rows[ns + ia[a]];
			}
#pragma unroll
			for(int b = 0; b < 3; b++){
				const double *rtb = rows + (2 + b) * ns, *r0b = rows + (5 + b) * ns;
				const double rt_jbi = rtb[i], r0_jbi = r0b[i];
#pragma unroll
				for(int a = 0; a < 3; a++){
					const double rt_jbia = rtb[ia[a]], r0_jbia = r0b[ia[a]];
					double corr = rt_ji * r0_jbia;
					corr += rt_jbia * r0_ji;
					corr -= rt_jbi * r0_jia[a];
					corr -= rt_jia[a] * r0_jbi;
					g[a * 3 + b] += chia[a] * chijb[b] * corr;
				}
			}
		}
	}
	// fixed-order sums: a butterfly within each wave, then the waves in order
	const unsigned lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
	for(int q = 0; q < 9; q++){
		double v = g[q];
#pragma unroll
		for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
		if(lane == 0) red[q][wv] = v;
	}
	__syncthreads();
	if(threadIdx.x < 9){
		double t = 0;
		for(int w = 0; w < CC_BT / 64; w++) t += red[threadIdx.x][w];
		partial[((size_t)tau * gridDim.x + blockIdx.x) * 9 + threadIdx.x] = t;
	}
}

// out[tau][q] = scale * sum of the block sums of slice tau: workgroup tau, wave q; lane l adds the
// blocks l, l + 64, ... and a butterfly adds the lanes (a fixed order)
__global__ __launch_bounds__(576) void cc_sum_kernel(const double *__restrict__ partial, unsigned nblocks, double scale,
		double *__restrict__ out)
{
	const unsigned tau = blockIdx.x, q = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const double *p = partial + (size_t)tau * nblocks * 9 + q;
	double s = 0;
	for(unsigned b = lane; b < nblocks; b += 64) s += p[(size_t)b * 9];
#pragma unroll
	for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
	if(lane == 0) out[tau * 9 + q] = s * scale;
}

// The correlator's kernels with the transfer-matrix stores Rt0, R0t (nt N^2 values each) and the
// block sums `partial` (cc_partials values) given: g_ab(tau) (nt x 9) into out, asynchronous.
static int correlate(ephmc_ctx *c, const double *x, const double *minv, double logdet, double *Rt0, double *R0t,
		double *partial, double *out)
{
	const unsigned ns = c->ns, nt = c->nt;
	EPHMC_TRY(coefficients(c, x, false));
	// ZERO_FILLING / SINGLE_PARTICLE start both propagators from the identity
	// (organic_dense_matrix.c:458-470)
	const bool ident = c->cfg.zero_filling || c->cfg.single_particle;
	// R_t0: slices nt-1 .. 0 applied to M_hat^{-1}, the state stored before every slice: step i
	// (slice nt-1-i) stores slot nt-i, so slot s >= 1 is stored entry nt-s; slot 0 is the result
	EPHMC_TRY(chain_forward(c, EPHMC_FP64, minv, c->d_Tfin, Rt0, 1, 1, ident));
	// R_0t: transposed inverse, then slices 0 .. nt-1 ascending; stored entry tau = state at tau
	if(!ident) EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, minv, c->d_gj[0], 0));
	EPHMC_TRY(chain_forward(c, EPHMC_FP64, c->d_gj[0], c->d_gj[1], R0t, 1, 0, ident));
	const unsigned by = (ns + CC_J - 1) / CC_J;
	const size_t lds = 8 * (size_t)ns * sizeof(double);
	if(lds + 9 * (CC_BT / 64) * sizeof(double) > 160 * 1024){ set_error("correlator: %u sites exceed the LDS", ns); return EPHMC_ERR_UNSUPPORTED; }
	(void)hipFuncSetAttribute((const void *)cc_contract_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	cc_contract_kernel<<<dim3(by, nt), CC_BT, lds, c->stream>>>(c->d_Tfin, Rt0, R0t, x, c->d_nb,
		c->cfg.t[0], c->cfg.t[1], c->cfg.t[2], c->cfg.kappa[0], c->cfg.kappa[1], c->cfg.kappa[2], ns, nt, partial);
	EPHMC_LAUNCHED();
	cc_sum_kernel<<<nt, 576, 0, c->stream>>>(partial, by, c->cfg.single_particle ? exp(-logdet) : 1.0 / ns,
		out);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// Allocation-free variant (ephmc_cc_correlator_static): its transfer-matrix stores and block sums
// live in this file-scope device array, 64 MiB on every GPU the library runs on, held until the
// process ends.  In host code `scratch` is only a symbol: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device on the entry point's first call there, and only that
// address is passed to kernels.  Calls that overlap on one GPU share (and race on) the array.
constexpr size_t CC_STATIC_DOUBLES = (size_t)64 << 17;     // 64 MiB of doubles
static __device__ __align__(256) double scratch[CC_STATIC_DOUBLES];

constexpr int CC_MAX_DEVICES = 64;
static double *g_scratch_addr[CC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of `scratch` on the context's GPU (looked up once per device)
static int static_scratch(ephmc_ctx *c, double **out)
{
	if(c->device < 0 || c->device >= CC_MAX_DEVICES){ set_error("device %d out of range", c->device); return EPHMC_ERR_ARG; }
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if(!g_scratch_addr[c->device]){
		int prev = 0;
		EPHMC_HIP(hipGetDevice(&prev));
		EPHMC_HIP(hipSetDevice(c->device));
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		(void)hipSetDevice(prev);
		if(e != hipSuccess || !p){
			(void)hipGetLastError();
			set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", c->device, hipGetErrorString(e));
			return EPHMC_ERR_HIP;
		}
		g_scratch_addr[c->device] = (double *)p;
	}
	*out = g_scratch_addr[c->device];
	return EPHMC_OK;
}

} // namespace ephmc

using namespace ephmc;

extern "C" int ephmc_cc_correlator_static(ephmc_ctx *c, const double *x, const double *minv, double logdet, double *greens_dev)
{
	if(!c || !x || !greens_dev || (!minv && !c->cfg.zero_filling && !c->cfg.single_particle)){
		set_error("null argument");
		return EPHMC_ERR_ARG;
	}
	const size_t store = (size_t)c->nt * c->mat, need = 2 * store + cc_partials(c->ns, c->nt);
	if(need > CC_STATIC_DOUBLES){
		set_error("ephmc_cc_correlator_static: needs %zu doubles (2 nt N^2 + block sums), the static scratch holds %zu "
			"(64 MiB); use ephmc_cc_correlator", need, CC_STATIC_DOUBLES);
		return EPHMC_ERR_ARG;
	}
	double *base = nullptr;
	EPHMC_TRY(static_scratch(c, &base));
	// every value of the stores and block sums is written before it is read: nothing to clear
	return correlate(c, x, minv, logdet, base, base + store, base + 2 * store, greens_dev);
}

extern "C" int ephmc_cc_correlator(ephmc_ctx *c, const double *x, const double *minv, double logdet, double *greens_host)
{
	if(!c || !x || !greens_host || (!minv && !c->cfg.zero_filling && !c->cfg.single_particle)){
		set_error("null argument");
		return EPHMC_ERR_ARG;
	}
	const unsigned ns = c->ns, nt = c->nt;
	const size_t mat = c->mat;
	if(!c->d_Rt0 || !c->d_R0t || !c->d_ccg){
		// all three or none: a failed allocation releases the others, so the next call retries
		// instead of running on a partial set
		const int rc = cc_alloc(c, nt * mat, nt * 9 + cc_partials(ns, nt));
		if(rc != EPHMC_OK) return rc;
	}
	EPHMC_TRY(timer_begin(c, "correlator"));
	EPHMC_TRY(correlate(c, x, minv, logdet, c->d_Rt0, c->d_R0t, c->d_ccg + (size_t)nt * 9, c->d_ccg));
	EPHMC_TRY(timer_end(c, "correlator"));
	EPHMC_HIP(hipMemcpyAsync(greens_host, c->d_ccg, nt * 9 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	return EPHMC_OK;
}
