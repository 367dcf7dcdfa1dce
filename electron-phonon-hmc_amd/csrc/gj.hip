// gj.hip -- log|det| and inverse of M_hat by blocked Gauss-Jordan elimination with partial
// pivoting (FP64), replacing LAPACK dgetrf + dgetri in logDetM_and_inv_M_hat
// (organic_dense_matrix.c:547-607).
//
// Row pivots are chosen exactly as in LU with partial pivoting (largest |a| among the rows not
// yet pivoted, first index on ties, like idamax), so the pivot values are LU's U_ii and
// log|det| = sum log|d_g|, sign = prod sign(d_g) * (-1)^(#row swaps) (organic_dense_matrix.c:582-588).
//
// Block step for columns K = [k0, k0+nb):
//   panel  (one workgroup, rows in registers): the nb unblocked in-place Gauss-Jordan steps on
//          the N x nb block column; yields E (the block's non-trivial transform columns, which
//          are also the final in-place contents of columns K), the pivots and the composed row
//          permutation src.
//   update (all other columns, out of place):  W'[i,c] = [i not in K] W[src i, c]
//          + sum_kk E[i,kk] W[src(k0+kk), c].
// After the last block the in-place result is A^{-1} with its columns permuted; undoing the
// swaps in reverse order (a gather by the composed permutation) gives the inverse.
#include "internal.h"

namespace ephmc {

template <int NB, int RT, int BT>
__global__ __launch_bounds__(BT) void gj_panel_kernel(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe, int *__restrict__ piv, double *__restrict__ diag, int *__restrict__ src)
{
	extern __shared__ int rowid[];             // n entries
	__shared__ double red_val[BT / 64];
	__shared__ int red_idx[BT / 64];
	__shared__ double rowA[NB], rowB[NB];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

	double P[RT][NB];
#pragma unroll
	for(int j = 0; j < RT; j++){
		const unsigned r = tid + j * BT;
#pragma unroll
		for(int kk = 0; kk < NB; kk++)
			P[j][kk] = (r < n && kk < (int)nbe) ? A[(size_t)(k0 + kk) * n + r] : 0.0;
	}
	for(unsigned r = tid; r < n; r += BT) rowid[r] = (int)r;
	__syncthreads();

	for(int jj = 0; jj < (int)nbe; jj++){
		const unsigned g = k0 + jj;
		// pivot search among rows >= g
		double best = -1.0;
		int bidx = 0x7fffffff;
#pragma unroll
		for(int j = 0; j < RT; j++){
			const unsigned r = tid + j * BT;
			if(r < n && r >= g){
				double a = 0;
#pragma unroll
				for(int kk = 0; kk < NB; kk++) if(kk == jj) a = fabs(P[j][kk]);
				if(a > best){ best = a; bidx = (int)r; }
			}
		}
#pragma unroll
		for(int off = 32; off >= 1; off >>= 1){
			const double ob = __shfl_xor(best, off, 64);
			const int oi = __shfl_xor(bidx, off, 64);
			if(ob > best || (ob == best && oi < bidx)){ best = ob; bidx = oi; }
		}
		if(lane == 0){ red_val[wave] = best; red_idx[wave] = bidx; }
		__syncthreads();
		best = red_val[0]; bidx = red_idx[0];
#pragma unroll
		for(int w = 1; w < BT / 64; w++){
			const double ob = red_val[w];
			const int oi = red_idx[w];
			if(ob > best || (ob == best && oi < bidx)){ best = ob; bidx = oi; }
		}
		const unsigned p = (bidx < (int)n) ? (unsigned)bidx : g;   // all-NaN column: no swap
		// publish pivot row (old row p) and, when swapping, old row g
#pragma unroll
		for(int j = 0; j < RT; j++){
			const unsigned r = tid + j * BT;
			if(r == p){
#pragma unroll
				for(int kk = 0; kk < NB; kk++) rowA[kk] = P[j][kk];
			}
			if(p != g && r == g){
#pragma unroll
				for(int kk = 0; kk < NB; kk++) rowB[kk] = P[j][kk];
			}
		}
		if(tid == 0){
			piv[g] = (int)p;
			if(p != g){ const int t = rowid[g]; rowid[g] = rowid[p]; rowid[p] = t; }
		}
		__syncthreads();
		const double *u = rowA;    // pivot row, read as LDS broadcasts
#pragma unroll
		for(int j = 0; j < RT; j++){
			const unsigned r = tid + j * BT;
			if(r == g){
#pragma unroll
				for(int kk = 0; kk < NB; kk++) P[j][kk] = u[kk];
			}else if(p != g && r == p){
#pragma unroll
				for(int kk = 0; kk < NB; kk++) P[j][kk] = rowB[kk];
			}
		}
		double d = 0;
#pragma unroll
		for(int kk = 0; kk < NB; kk++) if(kk == jj) d = u[kk];
		const double inv = 1.0 / d;
		if(tid == 0) diag[g] = d;
#pragma unroll
		for(int j = 0; j < RT; j++){
			const unsigned r = tid + j * BT;
			if(r >= n) continue;
			if(r == g){
#pragma unroll
				for(int kk = 0; kk < NB; kk++) P[j][kk] = (kk == jj) ? inv : u[kk] * inv;
			}else{
				double f = 0;
#pragma unroll
				for(int kk = 0; kk < NB; kk++) if(kk == jj) f = P[j][kk];
				const double fi = f * inv;
#pragma unroll
				for(int kk = 0; kk < NB; kk++) P[j][kk] = (kk == jj) ? -fi : P[j][kk] - fi * u[kk];
			}
		}
	}
	__syncthreads();
#pragma unroll
	for(int j = 0; j < RT; j++){
		const unsigned r = tid + j * BT;
		if(r < n)
#pragma unroll
			for(int kk = 0; kk < NB; kk++)
				if(kk < (int)nbe) An[(size_t)(k0 + kk) * n + r] = P[j][kk];
	}
	for(unsigned r = tid; r < n; r += BT) src[r] = rowid[r];
}

// Rank-nb update of every column outside K (64 x 64 tiles, 4 x 4 outputs per thread).
template <int NB>
__global__ __launch_bounds__(256) void gj_update_kernel(const double *__restrict__ A, double *__restrict__ An,
		const int *__restrict__ src, unsigned n, unsigned k0, unsigned nbe)
{
	__shared__ double Y[NB][64 + 1];      // pivot rows of this column tile
	__shared__ double E[NB][64 + 1];      // transform rows of this row tile (transposed)
	__shared__ int rsrc[64];
	const unsigned row0 = blockIdx.x * 64, col0 = blockIdx.y * 64;
	const int tid = threadIdx.x;
	if(col0 >= k0 && col0 + 64 <= k0 + nbe) return;   // tile entirely inside K
	for(int e = tid; e < NB * 64; e += 256){
		const int kk = e / 64, cc = e % 64;
		const unsigned col = col0 + cc;
		Y[kk][cc] = (kk < (int)nbe && col < n) ? A[(size_t)col * n + src[k0 + kk]] : 0.0;
		const unsigned row = row0 + cc;
		E[kk][cc] = (kk < (int)nbe && row < n) ? An[(size_t)(k0 + kk) * n + row] : 0.0;
	}
	if(tid < 64) rsrc[tid] = (row0 + tid < n) ? src[row0 + tid] : 0;
	__syncthreads();
	const int tr = tid & 15, tc = tid >> 4;           // rows tr + 16 i, columns tc + 16 j
	double acc[4][4];
#pragma unroll
	for(int i = 0; i < 4; i++){
		const unsigned row = row0 + tr + 16 * i;
		const bool inK = row >= k0 && row < k0 + nbe;
#pragma unroll
		for(int j = 0; j < 4; j++){
			const unsigned col = col0 + tc + 16 * j;
			acc[i][j] = (row < n && col < n && !inK) ? A[(size_t)col * n + rsrc[tr + 16 * i]] : 0.0;
		}
	}
#pragma unroll 8
	for(int kk = 0; kk < NB; kk++){
		double e[4], y[4];
#pragma unroll
		for(int i = 0; i < 4; i++) e[i] = E[kk][tr + 16 * i];
#pragma unroll
		for(int j = 0; j < 4; j++) y[j] = Y[kk][tc + 16 * j];
#pragma unroll
		for(int i = 0; i < 4; i++)
#pragma unroll
			for(int j = 0; j < 4; j++) acc[i][j] += e[i] * y[j];
	}
#pragma unroll
	for(int j = 0; j < 4; j++){
		const unsigned col = col0 + tc + 16 * j;
		if(col >= n || (col >= k0 && col < k0 + nbe)) continue;
#pragma unroll
		for(int i = 0; i < 4; i++){
			const unsigned row = row0 + tr + 16 * i;
			if(row < n) An[(size_t)col * n + row] = acc[i][j];
		}
	}
}

// log|det|, sign and the column permutation that undoes the pivoting.
__global__ __launch_bounds__(1024) void gj_finalize_kernel(const double *__restrict__ diag, const int *__restrict__ piv,
		unsigned n, int *__restrict__ colsrc, double *__restrict__ scal)
{
	extern __shared__ int perm[];
	__shared__ double part[1024];
	__shared__ int negs[1024];
	const int tid = threadIdx.x;
	double s = 0;
	int neg = 0;
	for(unsigned i = tid; i < n; i += 1024){
		const double d = diag[i];
		s += log(fabs(d));
		neg += (d < 0) + (piv[i] != (int)i);
	}
	part[tid] = s;
	negs[tid] = neg;
	for(unsigned i = tid; i < n; i += 1024) perm[i] = (int)i;
	__syncthreads();
	for(int off = 512; off > 0; off >>= 1){
		if(tid < off){ part[tid] += part[tid + off]; negs[tid] += negs[tid + off]; }
		__syncthreads();
	}
	if(tid == 0){
		scal[SCAL_LOGDET] = part[0];
		scal[SCAL_SIGN] = (negs[0] & 1) ? -1.0 : 1.0;
		// columns swapped back in reverse pivot order: final column c = in-place column perm[c]
		for(int g = (int)n - 1; g >= 0; g--){
			const int p = piv[g];
			if(p != g){ const int t = perm[g]; perm[g] = perm[p]; perm[p] = t; }
		}
	}
	__syncthreads();
	for(unsigned i = tid; i < n; i += 1024) colsrc[i] = perm[i];
}

__global__ void gather_columns_kernel(const double *__restrict__ W, double *__restrict__ out,
		const int *__restrict__ colsrc, unsigned n)
{
	const unsigned col = blockIdx.y;
	const size_t from = (size_t)colsrc[col] * n, to = (size_t)col * n;
	for(unsigned r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x)
		out[to + r] = W[from + r];
}

template <int NB, int RT, int BT>
static int launch_panel(ephmc_ctx *c, const double *A, double *An, unsigned k0, unsigned nbe)
{
	gj_panel_kernel<NB, RT, BT><<<1, BT, c->ns * sizeof(int), c->stream>>>(A, An, c->ns, k0, nbe, c->d_piv, c->d_diag, c->d_src);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// M (column-major, N x N, FP64) <- M^{-1}; log|det| and sign into the device scalars.
int gj_logdet_inverse(ephmc_ctx *c, double *M)
{
	const unsigned n = c->ns;
	// panel width: 32 columns for 256-thread panels, 16 for 1024-thread panels (the
	// register-resident panel row then stays within the 128 VGPRs of a 16-wave workgroup)
	int nb;
	if(n <= 256) nb = 32;
	else if(n <= 2048) nb = 16;
	else { set_error("gj: N = %u exceeds 2048", n); return EPHMC_ERR_UNSUPPORTED; }
	const double *cur = M;
	double *buf[2] = {c->d_gj[0], c->d_gj[1]};
	int which = 0;
	const dim3 ugrid((n + 63) / 64, (n + 63) / 64);
	for(unsigned k0 = 0; k0 < n; k0 += nb){
		const unsigned nbe = std::min<unsigned>(nb, n - k0);
		double *next = buf[which];
		if(n <= 256) EPHMC_TRY((launch_panel<32, 1, 256>(c, cur, next, k0, nbe)));
		else if(n <= 1024) EPHMC_TRY((launch_panel<16, 1, 1024>(c, cur, next, k0, nbe)));
		else EPHMC_TRY((launch_panel<16, 2, 1024>(c, cur, next, k0, nbe)));
		if(nb == 32) gj_update_kernel<32><<<ugrid, 256, 0, c->stream>>>(cur, next, c->d_src, n, k0, nbe);
		else gj_update_kernel<16><<<ugrid, 256, 0, c->stream>>>(cur, next, c->d_src, n, k0, nbe);
		EPHMC_LAUNCHED();
		cur = next;
		which ^= 1;
	}
	gj_finalize_kernel<<<1, 1024, n * sizeof(int), c->stream>>>(c->d_diag, c->d_piv, n, c->d_src, c->d_scal);
	EPHMC_LAUNCHED();
	gather_columns_kernel<<<dim3((n + 255) / 256, n), 256, 0, c->stream>>>(cur, M, c->d_src, n);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

} // namespace ephmc
