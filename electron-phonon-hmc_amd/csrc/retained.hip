// retained.hip -- the remaining dense-matrix entry points of the reference's host interface,
// computed on the GPU: the bare stencil A_tau on any number of columns, products of pair
// factors (1 + a_lin A + a_quad A^2) on any number of columns, the chain-rule matrix of one
// slice (refine_QM1P), the per-slice matrices QM1P of one or of every slice, the boson force and
// the classical-Fourier-acceleration drift.  The molecular-dynamics force of the engine does not
// use these (it runs the adjoint sweep); they serve callers of the reference's functions.
//
// Storage is the reference's: an N x N matrix is N contiguous chunks of N values, each chunk a
// column (organic_dense_matrix.c:209-216 applies B to every chunk).  A right multiplication by a
// slice polynomial is done as transpose, left multiplication, transpose (every factor is
// symmetric).
#include "internal.h"

namespace ephmc {

// out[c][i] = mu in[c][i] + sum_d coef[d][i] in[c][nb(i, d)] for ncols chunks of ns values, with
// the coefficients of d_coef64 slot 0 (slice_coefficients); out-of-place
__global__ __launch_bounds__(256) void stencil_cols_kernel(const double *__restrict__ coef, const int *__restrict__ nb,
		double mu, unsigned ns, size_t total, const double *__restrict__ in, double *__restrict__ out)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x){
		const size_t col = e / ns;
		const unsigned i = (unsigned)(e - col * ns);
		const double *v = in + col * ns;
		double w = mu * v[i];
#pragma unroll
		for(int d = 0; d < 6; d++) w = fmac(coef[(size_t)d * ns + i], v[nb[(size_t)d * ns + i]], w);
		out[e] = w;
	}
}

// y <- y + (a u + b z) (z may be null: y <- y + a u), the reference's v += a_lin w + a_quad z order
__global__ __launch_bounds__(256) void combine_kernel(double *__restrict__ y, double a, const double *__restrict__ u,
		double b, const double *__restrict__ z, size_t n)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
		y[e] += z ? a * u[e] + b * z[e] : a * u[e];
}

// y <- y + a (u + v), refine_QM1P's QM1P += a_quad (loc + loc2)
__global__ __launch_bounds__(256) void sum_into_kernel(double *__restrict__ y, double a, const double *__restrict__ u,
		const double *__restrict__ v, size_t n)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
		y[e] += a * (u[e] + v[e]);
}

// a fill kernel rather than hipMemsetAsync, so that callers may capture these calls in a graph
__global__ __launch_bounds__(256) void zero_cols_kernel(double *__restrict__ y, size_t n)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) y[e] = 0;
}

static unsigned grid_for(size_t n){ return (unsigned)std::min<size_t>((n + 255) / 256, 8192); }

// out = A_tau in over ncols chunks (coefficients in slot 0)
static int stencil_cols(ephmc_ctx *c, const double *in, double *out, size_t ncols)
{
	const size_t n = ncols * c->ns;
	if(n == 0) return EPHMC_OK;
	stencil_cols_kernel<<<grid_for(n), 256, 0, c->stream>>>(c->d_coef64, c->d_nb, c->cfg.mu, c->ns, n, in, out);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// M <- prod_j (1 + a_lin[j] A + a_quad[j] A^2) M over ncols chunks, j in the given order, with the
// slice-0 coefficients; t1, t2: scratch of ncols chunks each
static int pairs_cols(ephmc_ctx *c, double *M, size_t ncols, int npairs, const double *a_lin, const double *a_quad,
		double *t1, double *t2)
{
	const size_t n = ncols * c->ns;
	for(int j = 0; j < npairs; j++){
		EPHMC_TRY(stencil_cols(c, M, t1, ncols));
		const bool quad = a_quad[j] != 0;
		if(quad) EPHMC_TRY(stencil_cols(c, t1, t2, ncols));
		combine_kernel<<<grid_for(n), 256, 0, c->stream>>>(M, a_lin[j], t1, a_quad[j], quad ? t2 : nullptr, n);
		EPHMC_LAUNCHED();
	}
	return EPHMC_OK;
}

// the context's pair factors selected as in apply_pair_expansion_mu_K_to_v
// (organic_dense_matrix.c:186-194): positions [0, stop) forward, or np-1 down to stop+1
static int select_pairs(const ephmc_ctx *c, int revert, int stop, double *a_lin, double *a_quad)
{
	int n = 0;
	auto add = [&](int p){ a_lin[n] = c->poly.a_lin[p]; a_quad[n] = c->poly.cplx[p] ? c->poly.a_quad[p] : 0.0; n++; };
	if(revert){ for(int p = c->poly.np - 1; p > stop; p--) add(p); }
	else{ for(int p = 0; p < stop && p < c->poly.np; p++) add(p); }
	return n;
}

static int copy_d2d(ephmc_ctx *c, double *dst, const double *src, size_t n)
{
	if(dst != src) EPHMC_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	return EPHMC_OK;
}

// Q <- sum_j a_lin_j Z_j + a_quad_j (A Z_j + Z_j A), Z_j = P_<j Q P_>j (P: the pair factors
// before / after position j): the chain rule through the pair product, so that
// tr[Q dB] = tr[Y dA] (refine_QM1P, organic_dense_matrix.c:383-442).  x_tau: 3N couplings.
// Scratch: d_gj[0] (copy of Q), d_gj[1] (Z_j), d_Tfin, d_mhat.
static int refine(ephmc_ctx *c, const double *x_tau, double *Q)
{
	const size_t mat = c->mat;
	double *cp = c->d_gj[0], *z = c->d_gj[1], *t1 = c->d_Tfin, *t2 = c->d_mhat;
	EPHMC_TRY(copy_d2d(c, cp, Q, mat));
	zero_cols_kernel<<<grid_for(mat), 256, 0, c->stream>>>(Q, mat);
	EPHMC_LAUNCHED();
	for(int j = 0; j < c->poly.np; j++){
		EPHMC_TRY(copy_d2d(c, z, cp, mat));
		EPHMC_TRY(apply_exp_partial(c, x_tau, z, 0, j));                 // P_<j Q
		EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, z, t1, 0));
		EPHMC_TRY(apply_exp_partial(c, x_tau, t1, 1, j));                // (P_<j Q P_>j)^T
		EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, t1, z, 0));     // Z_j
		combine_kernel<<<grid_for(mat), 256, 0, c->stream>>>(Q, c->poly.a_lin[j], z, 0.0, nullptr, mat);
		EPHMC_LAUNCHED();
		if(!c->poly.cplx[j]) continue;
		// slot 0 holds this slice's coefficients (apply_exp_partial wrote them)
		EPHMC_TRY(stencil_cols(c, z, t1, c->ns));                         // A Z_j
		EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, z, t2, 0));     // Z_j^T
		EPHMC_TRY(stencil_cols(c, t2, z, c->ns));                         // A Z_j^T
		EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, z, t2, 0));     // Z_j A
		sum_into_kernel<<<grid_for(mat), 256, 0, c->stream>>>(Q, c->poly.a_quad[j], t1, t2, mat);
		EPHMC_LAUNCHED();
	}
	return EPHMC_OK;
}

// Q <- Q B_tau0 B_tau0+1 ... (ntau slices ascending) by transposes around a left chain; t1, t2
// scratch
static int right_slices(ephmc_ctx *c, double *Q, int tau0, int ntau, double *t1, double *t2)
{
	if(ntau <= 0) return EPHMC_OK;
	EPHMC_TRY(transpose_plus_identity(c, EPHMC_FP64, Q, t1, 0));
	EPHMC_TRY(chain_range(c, EPHMC_FP64, t1, t2, tau0, 1, ntau));
	return transpose_plus_identity(c, EPHMC_FP64, t2, Q, 0);
}

// Every QM1P_tau = B_{tau+1} ... B_{nt-1} G B_0 ... B_{tau-1} (G = M_hat^{-1}, or 1 for
// SINGLE_PARTICLE) by recursive halving: slot a holds S(a, b) = B_b ... B_{nt-1} G B_0 ... B_{a-1}
// for the slot range [a, b); splitting at m gives S(m, b) = S(a, b) B_a ... B_{m-1} (into slot m)
// and S(a, m) = B_m ... B_{b-1} S(a, b) (in place), down to ranges of one slot, where
// S(tau, tau+1) = QM1P_tau: nt log2(nt) slice applications, as the reference's binary splitting
// (construct_all_QM1P, organic_dense_matrix.c:311-381).  Coefficients of all slices must be current.
static int all_qm1p(ephmc_ctx *c, double *Q, int a, int b)
{
	if(b - a <= 1) return EPHMC_OK;
	const int m = a + (b - a + 1) / 2;
	const size_t mat = c->mat;
	double *sa = Q + (size_t)a * mat, *sm = Q + (size_t)m * mat;
	EPHMC_TRY(copy_d2d(c, sm, sa, mat));
	EPHMC_TRY(right_slices(c, sm, a, m - a, c->d_gj[0], c->d_gj[1]));
	EPHMC_TRY(chain_range(c, EPHMC_FP64, sa, sa, b - 1, -1, b - m));
	EPHMC_TRY(all_qm1p(c, Q, a, m));
	return all_qm1p(c, Q, m, b);
}

} // namespace ephmc

using namespace ephmc;

extern "C" {

int ephmc_apply_mu_K(ephmc_ctx *c, const double *x_tau, const double *in, double *out, unsigned ncols)
{
	if(!c || !x_tau || !in || !out){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(in == out){ set_error("ephmc_apply_mu_K is out of place: in and out must differ"); return EPHMC_ERR_ARG; }
	if(ncols > c->ns){ set_error("ephmc_apply_mu_K: %u columns exceed N = %u", ncols, c->ns); return EPHMC_ERR_ARG; }
	EPHMC_TRY(slice_coefficients(c, x_tau));
	return stencil_cols(c, in, out, ncols);
}

int ephmc_apply_pairs(ephmc_ctx *c, const double *x_tau, double *M, unsigned ncols, int npairs, const double *a_lin,
		const double *a_quad)
{
	if(!c || !x_tau || !M || (npairs > 0 && (!a_lin || !a_quad))){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(ncols > c->ns){ set_error("ephmc_apply_pairs: %u columns exceed N = %u", ncols, c->ns); return EPHMC_ERR_ARG; }
	if(npairs <= 0 || ncols == 0) return EPHMC_OK;
	EPHMC_TRY(slice_coefficients(c, x_tau));
	// scratch: two N x N buffers hold any ncols <= N chunks
	return pairs_cols(c, M, ncols, npairs, a_lin, a_quad, c->d_gj[0], c->d_gj[1]);
}

int ephmc_apply_exp_cols(ephmc_ctx *c, const double *x_tau, double *M, unsigned ncols, int revert, int stop)
{
	if(!c || !x_tau || !M){ set_error("null argument"); return EPHMC_ERR_ARG; }
	double al[MAX_PAIRS], aq[MAX_PAIRS];
	const int n = select_pairs(c, revert, stop, al, aq);
	return ephmc_apply_pairs(c, x_tau, M, ncols, n, al, aq);
}

int ephmc_refine_QM1P(ephmc_ctx *c, const double *x_tau, double *Q)
{
	if(!c || !x_tau || !Q){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return refine(c, x_tau, Q);
}

int ephmc_construct_QM1P(ephmc_ctx *c, const double *x, const double *M_inv, double *Q, int time)
{
	const bool sp = c && c->cfg.single_particle;
	if(!c || !x || !Q || (!M_inv && !sp)){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(time < 0 || time >= (int)c->nt){ set_error("construct_QM1P: time %d outside [0, %u)", time, c->nt); return EPHMC_ERR_ARG; }
	EPHMC_TRY(coefficients(c, x, false));
	// B_{time+1} ... B_{nt-1} G (slices nt-1 down to time+1; none leaves a copy of G)
	const int nleft = (int)c->nt - 1 - time;
	if(nleft > 0) EPHMC_TRY(chain_range(c, EPHMC_FP64, sp ? nullptr : M_inv, Q, (int)c->nt - 1, -1, nleft));
	else if(sp) EPHMC_TRY(set_identity(c, EPHMC_FP64, Q));
	else EPHMC_TRY(copy_d2d(c, Q, M_inv, c->mat));
	EPHMC_TRY(right_slices(c, Q, 0, time, c->d_gj[0], c->d_gj[1]));
	return refine(c, x + (size_t)time * c->loc, Q);
}

int ephmc_construct_all_QM1P(ephmc_ctx *c, const double *x, const double *M_inv, double *Q)
{
	const bool sp = c && c->cfg.single_particle;
	if(!c || !x || !Q || (!M_inv && !sp)){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_TRY(coefficients(c, x, false));
	if(sp) EPHMC_TRY(set_identity(c, EPHMC_FP64, Q));
	else EPHMC_TRY(copy_d2d(c, Q, M_inv, c->mat));
	return all_qm1p(c, Q, 0, (int)c->nt);
}

int ephmc_boson_force(ephmc_ctx *c, const double *x, double *p_dot, int overwrite)
{
	if(!c || !x || !p_dot){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return boson_force(c, x, p_dot, overwrite);
}

int ephmc_evolve_bosons_approx_step(ephmc_ctx *c, double *x, const double *p, double h)
{
	if(!c || !x || !p){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return fa_drift(c, p, x, h);
}

} // extern "C"
