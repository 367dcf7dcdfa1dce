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
// The (i, j) sum is split over workgroups (partials per workgroup, then a fixed-order pass).
#include "internal.h"

namespace ephmc {

constexpr int CC_BT = 256;
constexpr int CC_ROWS = 8;     // values of j per thread

__global__ __launch_bounds__(CC_BT) void cc_contract_kernel(const double *__restrict__ Rt0, const double *__restrict__ R0t,
		const double *__restrict__ x, const int *__restrict__ nb, double t0, double t1, double t2,
		double k0, double k1, double k2, unsigned ns, unsigned tau, double *__restrict__ partial)
{
	__shared__ double red[9][CC_BT];
	const double tk[3] = {t0, t1, t2}, kk[3] = {k0, k1, k2};
	double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
	const unsigned i = blockIdx.y;                      // one i per grid row
	const double *xt = x + (size_t)tau * 3 * ns;
	unsigned ia[3];
	double chia[3];
#pragma unroll
	for(int a = 0; a < 3; a++){
		ia[a] = (unsigned)nb[(2 * a) * ns + i];
		chia[a] = tk[a] - kk[a] * x[(size_t)i * 3 + a];
	}
	for(unsigned jj = 0; jj < CC_ROWS; jj++){
		const unsigned j = (blockIdx.x * CC_ROWS + jj) * CC_BT + threadIdx.x;
		if(j >= ns) break;
		const double rt_ji = Rt0[(size_t)j * ns + i], r0_ji = R0t[(size_t)j * ns + i];
		double rt_jia[3], r0_jia[3];
#pragma unroll
		for(int a = 0; a < 3; a++){
			rt_jia[a] = Rt0[(size_t)j * ns + ia[a]];
			r0_jia[a] = R0t[(size_t)j * ns + ia[a]];
		}
#pragma unroll
		for(int b = 0; b < 3; b++){
			const unsigned jb = (unsigned)nb[(2 * b) * ns + j];
			const double chijb = tk[b] - kk[b] * xt[(size_t)j * 3 + b];
			const double rt_jbi = Rt0[(size_t)jb * ns + i], r0_jbi = R0t[(size_t)jb * ns + i];
#pragma unroll
			for(int a = 0; a < 3; a++){
				const double rt_jbia = Rt0[(size_t)jb * ns + ia[a]], r0_jbia = R0t[(size_t)jb * ns + ia[a]];
				double corr = rt_ji * r0_jbia;
				corr += rt_jbia * r0_ji;
				corr -= rt_jbi * r0_jia[a];
				corr -= rt_jia[a] * r0_jbi;
				g[a * 3 + b] += chia[a] * chijb * corr;
			}
		}
	}
#pragma unroll
	for(int q = 0; q < 9; q++) red[q][threadIdx.x] = g[q];
	__syncthreads();
	for(int off = CC_BT / 2; off > 0; off >>= 1){
		if((int)threadIdx.x < off)
#pragma unroll
			for(int q = 0; q < 9; q++) red[q][threadIdx.x] += red[q][threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x < 9) partial[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 9 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void cc_sum_kernel(const double *__restrict__ partial, unsigned nblocks, double scale, double *__restrict__ out)
{
	const unsigned q = threadIdx.x;
	if(q >= 9) return;
	double s = 0;
	for(unsigned b = 0; b < nblocks; b++) s += partial[(size_t)b * 9 + q];
	out[q] = s * scale;
}

} // namespace ephmc

using namespace ephmc;

extern "C" int ephmc_cc_correlator(ephmc_ctx *c, const double *x, const double *minv, double logdet, double *greens_host)
{
	if(!c || !x || !greens_host || (!minv && !c->cfg.zero_filling && !c->cfg.single_particle)){
		set_error("null argument");
		return EPHMC_ERR_ARG;
	}
	const unsigned ns = c->ns, nt = c->nt;
	const size_t mat = c->mat;
	if(!c->d_Rt0){
		if(hipMalloc((void **)&c->d_Rt0, nt * mat * sizeof(double)) != hipSuccess
			|| hipMalloc((void **)&c->d_R0t, nt * mat * sizeof(double)) != hipSuccess
			|| hipMalloc((void **)&c->d_ccg, nt * 9 * sizeof(double)) != hipSuccess){
			(void)hipGetLastError();
			set_error("correlator buffers (%zu MB) could not be allocated", 2 * nt * mat * 8 >> 20);
			return EPHMC_ERR_ALLOC;
		}
	}
	EPHMC_TRY(timer_begin(c, "correlator"));
	EPHMC_TRY(coefficients(c, x, false));
	// ZERO_FILLING / SINGLE_PARTICLE start both propagators from the identity
	// (organic_dense_matrix.c:458-470)
	const bool ident = c->cfg.zero_filling || c->cfg.single_particle;
	// R_t0: slices nt-1 .. 0 applied to M_hat^{-1}, the state stored before every slice: step i
	// (slice nt-1-i) stores slot nt-i, so slot s >= 1 is stored entry nt-s; slot 0 is the result
	EPHMC_TRY(chain_forward(c, EPHMC_FP64, minv, c->d_Tfin, c->d_Rt0, 1, 1, ident));
	// R_0t: transposed inverse, then slices 0 .. nt-1 ascending; stored entry tau = state at tau
	if(!ident) EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, minv, c->d_gj[0], 0));
	EPHMC_TRY(chain_forward(c, EPHMC_FP64, c->d_gj[0], c->d_gj[1], c->d_R0t, 1, 0, ident));
	const unsigned bx = (ns + CC_BT * CC_ROWS - 1) / (CC_BT * CC_ROWS);
	const dim3 grid(bx, ns);
	double *partial = c->d_gj[1];                   // ns * bx * 9 doubles <= N^2
	for(unsigned tau = 0; tau < nt; tau++){
		const double *rt = (tau == 0) ? c->d_Tfin : c->d_Rt0 + (size_t)(nt - tau) * mat;
		cc_contract_kernel<<<grid, CC_BT, 0, c->stream>>>(rt, c->d_R0t + (size_t)tau * mat, x, c->d_nb,
			c->cfg.t[0], c->cfg.t[1], c->cfg.t[2], c->cfg.kappa[0], c->cfg.kappa[1], c->cfg.kappa[2], ns, tau, partial);
		EPHMC_LAUNCHED();
		cc_sum_kernel<<<1, 64, 0, c->stream>>>(partial, bx * ns, c->cfg.single_particle ? exp(-logdet) : 1.0 / ns,
			c->d_ccg + (size_t)tau * 9);
		EPHMC_LAUNCHED();
	}
	EPHMC_TRY(timer_end(c, "correlator"));
	EPHMC_HIP(hipMemcpyAsync(greens_host, c->d_ccg, nt * 9 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	return EPHMC_OK;
}
