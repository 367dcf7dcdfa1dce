// gj.hip -- log|det| and inverse of M_hat (FP64), replacing LAPACK dgetrf + dgetri in
// logDetM_and_inv_M_hat (organic_dense_matrix.c:547-607).
//
// LAPACK's partial pivoting picks, at elimination step g, the row with the largest |a_rg|
// among rows r >= g (first index on ties).  Elimination without row exchanges makes the same
// choices exactly when every multiplier |l_rg| = |a_rg| / |a_gg| is <= 1; the pivots are then
// the same U_ii, so log|det| = sum log|U_ii| and the sign (organic_dense_matrix.c:582-588)
// agree with LAPACK, and the inverse is the same matrix up to rounding.  For M_hat = 1 + B..B at
// the lattice parameters of the reference (mu at the band bottom) the multipliers are of order
// 1e-2, so:
//
//   fast path: blocked Gauss-Jordan without row exchanges, 32 columns per block K = [k0, k0+nb):
//     panel  D^{-1} of the diagonal block D = A[K,K] (one wave), then every row i:
//            E[i,K] = -A[i,K] D^{-1} (i not in K) or D^{-1} rows (i in K); the LU multipliers
//            of D and of the rows below K are checked |l| <= 1
//     update every column c outside K: A'[:,c] = [i not in K] A[:,c] + E[:,K] A[K,c]
//   check:   any multiplier > 1 sets a device flag; the single-workgroup fallback kernel, always
//            launched, then inverts the untouched input with partial pivoting (in-place
//            Gauss-Jordan with row exchanges and the columns un-permuted at the end); without
//            the flag it returns at once.  No host round trip either way.
#include "internal.h"
#include <cstdlib>

namespace ephmc {

constexpr int GJ_NB = 32;

__global__ void gj_clear_kernel(double *scal)
{
	scal[SCAL_GJ_FLAG] = 0;
}

// Block step, part 1 (grid over rows, 256 threads per workgroup).
// Every workgroup first inverts the diagonal block D = A[K,K] by Gauss-Jordan without row
// exchanges: thread t owns the four entries (r, 4q..4q+3), r = t/8, q = t%8, in registers; at
// step g the owners of row g and of column g publish them to LDS (double-buffered by step
// parity, one barrier per step).  D^{-1} and the LU multipliers (rows below the pivot) end in
// LDS; workgroup 0 records the pivots and checks |L| <= 1.  Then every thread forms the
// transform row of one matrix row i:
//   i in K:      E[i,:] = D^{-1}[i-k0,:]
//   i not in K:  E[i,:] = -A[i,K] D^{-1};  rows below K check their LU multipliers
//                l = A[i,K] U^{-1} = -E[i,:] L  (|l| <= 1).
template <int NB>
__global__ __launch_bounds__(256) void gj_panel_np_kernel(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe, double *__restrict__ diag, int *__restrict__ piv, double *__restrict__ scal)
{
	static_assert(NB == 32, "one 32 x 32 block per 256 threads");
	__shared__ double dinv[NB][NB + 1];
	__shared__ double lm[NB][NB + 1];        // lm[r][g]: multiplier of row r at step g (r > g)
	__shared__ double rowb[2][NB], colb[2][NB];
	const int tid = threadIdx.x;
	const int r = tid >> 3, c0 = (tid & 7) * 4;
	double v[4];
#pragma unroll
	for(int q = 0; q < 4; q++){
		const int cc = c0 + q;
		v[q] = (r < (int)nbe && cc < (int)nbe) ? A[(size_t)(k0 + cc) * n + k0 + r] : (r == cc ? 1.0 : 0.0);
		lm[r][cc] = 0;
	}
	bool bad = false;
	for(int g = 0; g < (int)nbe; g++){
		const int b = g & 1;
		if(r == g){
#pragma unroll
			for(int q = 0; q < 4; q++) rowb[b][c0 + q] = v[q];
		}
		if(g >= c0 && g < c0 + 4){
#pragma unroll
			for(int q = 0; q < 4; q++) if(c0 + q == g) colb[b][r] = v[q];
		}
		__syncthreads();
		const double d = rowb[b][g];
		const double inv = 1.0 / d;
		if(r == g){
#pragma unroll
			for(int q = 0; q < 4; q++) v[q] = (c0 + q == g) ? inv : v[q] * inv;
		}else{
			const double fi = colb[b][r] * inv;
#pragma unroll
			for(int q = 0; q < 4; q++) v[q] = (c0 + q == g) ? -fi : fmac(-fi, rowb[b][c0 + q], v[q]);
			if(r > g && g >= c0 && g < c0 + 4){
				lm[r][g] = fi;
				bad |= !(fabs(fi) <= 1.0);
			}
		}
		if(tid == 0 && blockIdx.x == 0){ diag[k0 + g] = d; piv[k0 + g] = (int)(k0 + g); }
	}
#pragma unroll
	for(int q = 0; q < 4; q++) dinv[r][c0 + q] = v[q];
	if(bad && blockIdx.x == 0) scal[SCAL_GJ_FLAG] = 1;
	__syncthreads();
	const unsigned i = blockIdx.x * 256 + tid;
	if(i >= n) return;
	const bool inK = i >= k0 && i < k0 + nbe;
	double e[NB];
	if(inK){
#pragma unroll
		for(int j = 0; j < NB; j++) e[j] = dinv[i - k0][j];
	}else{
		double a[NB];
#pragma unroll
		for(int m = 0; m < NB; m++) a[m] = m < (int)nbe ? A[(size_t)(k0 + m) * n + i] : 0.0;
#pragma unroll
		for(int j = 0; j < NB; j++){
			double s = 0;
#pragma unroll
			for(int m = 0; m < NB; m++) s = fmac(a[m], dinv[m][j], s);
			e[j] = -s;
		}
		if(i >= k0 + nbe){
			bool bad = false;
#pragma unroll
			for(int j = 0; j < NB; j++){
				double y = e[j];
#pragma unroll
				for(int m = j + 1; m < NB; m++) y += e[m] * lm[m][j];
				bad |= !(fabs(y) <= 1.0);
			}
			if(bad) scal[SCAL_GJ_FLAG] = 1;
		}
	}
#pragma unroll
	for(int j = 0; j < NB; j++)
		if(j < (int)nbe) An[(size_t)(k0 + j) * n + i] = e[j];
}

// Block step, part 2: rank-nb update of every column outside K (32 x 32 tiles, 2 x 2 per thread).
template <int NB>
__global__ __launch_bounds__(256) void gj_update_kernel(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe)
{
	__shared__ double Y[NB][32 + 1];      // rows K of this column tile
	__shared__ double E[NB][32 + 1];      // transform columns of this row tile (transposed)
	const unsigned row0 = blockIdx.x * 32, col0 = blockIdx.y * 32;
	const int tid = threadIdx.x;
	if(col0 >= k0 && col0 + 32 <= k0 + nbe) return;   // tile entirely inside K
	for(int q = tid; q < NB * 32; q += 256){
		const int kk = q / 32, cc = q % 32;
		const unsigned col = col0 + cc;
		Y[kk][cc] = (kk < (int)nbe && col < n) ? A[(size_t)col * n + k0 + kk] : 0.0;
		const unsigned row = row0 + cc;
		E[kk][cc] = (kk < (int)nbe && row < n) ? An[(size_t)(k0 + kk) * n + row] : 0.0;
	}
	__syncthreads();
	const int tr = tid & 31, tc = tid >> 5;           // row tr, columns tc + 8 j
	const unsigned row = row0 + tr;
	const bool rowK = row >= k0 && row < k0 + nbe;
	double acc[4];
#pragma unroll
	for(int j = 0; j < 4; j++){
		const unsigned col = col0 + tc + 8 * j;
		acc[j] = (row < n && col < n && !rowK) ? A[(size_t)col * n + row] : 0.0;
	}
#pragma unroll 8
	for(int kk = 0; kk < NB; kk++){
		const double ev = E[kk][tr];
#pragma unroll
		for(int j = 0; j < 4; j++) acc[j] = fmac(ev, Y[kk][tc + 8 * j], acc[j]);
	}
	if(row >= n) return;
#pragma unroll
	for(int j = 0; j < 4; j++){
		const unsigned col = col0 + tc + 8 * j;
		if(col >= n || (col >= k0 && col < k0 + nbe)) continue;
		An[(size_t)col * n + row] = acc[j];
	}
}

// Fused block step: one launch per 32-column block K, grid over 64 x 64 output tiles.
// Every workgroup inverts D = A[K,K] in LDS (the same arithmetic as gj_panel_np_kernel), forms
// the transform rows E[i,:] of its 64 rows (D^{-1} rows inside K, -A[i,K] D^{-1} outside), and
// writes its tile of A' = [i not in K] A + E A[K,:] (columns inside K: E itself).  Workgroups
// of the first column block check the LU multipliers of their rows below K; workgroup (0,0)
// records the pivots and the multipliers inside D.
constexpr int GJ_RB = 64, GJ_CB = 64;

__global__ __launch_bounds__(256) void gj_step_kernel(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe, double *__restrict__ diag, int *__restrict__ piv, double *__restrict__ scal)
{
	constexpr int NB = GJ_NB;
	__shared__ double dinv[NB][NB + 1];
	__shared__ double lm[NB][NB + 1];
	__shared__ double colb[2][NB];
	__shared__ double Es[GJ_RB][NB + 1];
	__shared__ double Ys[NB][GJ_CB + 1];
	__shared__ double Ar[GJ_RB][NB + 1];     // A[row block, K]
	const int tid = threadIdx.x;
	const unsigned r0 = blockIdx.x * GJ_RB, c0 = blockIdx.y * GJ_CB;
	const bool lead = blockIdx.x == 0 && blockIdx.y == 0;
	const int wave = tid >> 6, lane = tid & 63;
	if(wave == 0){
		// ---- D^{-1} by one wave, no barriers: lane c holds column c of [D | 1] (lanes 0..31 D,
		// 32..63 the identity); step g broadcasts the pivot column through LDS (colb).  After
		// the last step lanes 32..63 hold the columns of D^{-1}.
		double col[NB];
#pragma unroll
		for(int r = 0; r < NB; r++){
			if(lane < NB) col[r] = (r < (int)nbe && lane < (int)nbe) ? A[(size_t)(k0 + lane) * n + k0 + r] : (r == lane ? 1.0 : 0.0);
			else col[r] = (r == lane - NB) ? 1.0 : 0.0;
			if(lane < NB) lm[r][lane] = 0;
		}
		bool bad = false;
#pragma unroll
		for(int g = 0; g < NB; g++){
			if(g < (int)nbe){
				double *pc = colb[g & 1];
				if(lane == g){
#pragma unroll
					for(int r = 0; r < NB; r++) pc[r] = col[r];
				}
				__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
				__builtin_amdgcn_wave_barrier();
				const double d = pc[g];
				const double inv = 1.0 / d;
				const double pr = col[g];
#pragma unroll
				for(int r = 0; r < NB; r++){
					if(r == g) col[r] = pr * inv;
					else{
						const double fi = pc[r] * inv;
						col[r] = fmac(-fi, pr, col[r]);
						if(r > g){
							if(lane == 0) lm[r][g] = fi;
							bad |= !(fabs(fi) <= 1.0);
						}
					}
				}
				if(lane == 0 && lead){ diag[k0 + g] = d; piv[k0 + g] = (int)(k0 + g); }
				__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
				__builtin_amdgcn_wave_barrier();
			}
		}
		if(lane >= NB){
#pragma unroll
			for(int r = 0; r < NB; r++) dinv[r][lane - NB] = col[r];
		}
		if(bad && lead && lane == 0) scal[SCAL_GJ_FLAG] = 1;
	}
	// ---- loads that do not depend on D^{-1}, issued by waves 1..3 while wave 0 inverts:
	// A[K, column block] (old values; the output goes to the other buffer) and A[row block, K]
	for(int e = tid - 64; e >= 0 && e < NB * GJ_CB; e += 192){
		const int m = e % NB, cc = e / NB;
		const unsigned col = c0 + cc;
		Ys[m][cc] = (m < (int)nbe && col < n) ? A[(size_t)col * n + k0 + m] : 0.0;
	}
	for(int e = tid - 64; e >= 0 && e < GJ_RB * NB; e += 192){
		const int rr = e % GJ_RB, m = e / GJ_RB;
		const unsigned i = r0 + rr;
		Ar[rr][m] = (m < (int)nbe && i < n) ? A[(size_t)(k0 + m) * n + i] : 0.0;
	}
	// this thread's tile values: row tid % 64, columns (tid / 64) + 4 j
	const int trr = tid & 63, cbase = tid >> 6;
	const unsigned ti = r0 + trr;
	const bool rowK = ti >= k0 && ti < k0 + nbe;
	double av[GJ_CB / 4];
#pragma unroll
	for(int j = 0; j < GJ_CB / 4; j++){
		const unsigned col = c0 + cbase + 4 * j;
		av[j] = (ti < n && col < n && !rowK) ? A[(size_t)col * n + ti] : 0.0;
	}
	__syncthreads();
	// ---- transform rows of this row block: thread (row tid/4, 8 columns)
	{
		const int rr = tid >> 2, j0 = (tid & 3) * 8;
		const unsigned i = r0 + rr;
		double e8[8];
		if(i < n && i >= k0 && i < k0 + nbe){
#pragma unroll
			for(int j = 0; j < 8; j++) e8[j] = dinv[i - k0][j0 + j];
		}else{
#pragma unroll
			for(int j = 0; j < 8; j++) e8[j] = 0;
#pragma unroll 8
			for(int m = 0; m < NB; m++){
				const double a = Ar[rr][m];
#pragma unroll
				for(int j = 0; j < 8; j++) e8[j] = fmac(a, dinv[m][j0 + j], e8[j]);
			}
#pragma unroll
			for(int j = 0; j < 8; j++) e8[j] = -e8[j];
		}
#pragma unroll
		for(int j = 0; j < 8; j++) Es[rr][j0 + j] = e8[j];
	}
	__syncthreads();
	// LU multipliers of the rows below K: l = -E L (first column block only)
	if(blockIdx.y == 0 && tid < GJ_RB){
		const unsigned i = r0 + tid;
		if(i < n && i >= k0 + nbe){
			bool bad = false;
			for(int j = 0; j < (int)nbe; j++){
				double y = Es[tid][j];
				for(int m = j + 1; m < (int)nbe; m++) y += Es[tid][m] * lm[m][j];
				bad |= !(fabs(y) <= 1.0);
			}
			if(bad) scal[SCAL_GJ_FLAG] = 1;
		}
	}
	// ---- tile update
	if(ti >= n) return;
	double er[NB];
#pragma unroll
	for(int m = 0; m < NB; m++) er[m] = Es[trr][m];
#pragma unroll
	for(int j = 0; j < GJ_CB / 4; j++){
		const int cc = cbase + 4 * j;
		const unsigned col = c0 + cc;
		if(col >= n) continue;
		double out;
		if(col >= k0 && col < k0 + nbe) out = Es[trr][col - k0];
		else{
			double acc = av[j];
#pragma unroll
			for(int m = 0; m < NB; m++) acc = fmac(er[m], Ys[m][cc], acc);
			out = acc;
		}
		An[(size_t)col * n + ti] = out;
	}
}

// log|det| and sign from the pivots of the fast path (no row exchanges).
__global__ __launch_bounds__(1024) void gj_logdet_kernel(const double *__restrict__ diag, unsigned n, double *__restrict__ scal)
{
	__shared__ double part[1024];
	__shared__ int negs[1024];
	const int tid = threadIdx.x;
	if(scal[SCAL_GJ_FLAG] != 0) return;
	double s = 0;
	int neg = 0;
	for(unsigned i = tid; i < n; i += 1024){
		const double d = diag[i];
		s += log(fabs(d));
		neg += d < 0;
	}
	part[tid] = s;
	negs[tid] = neg;
	__syncthreads();
	for(int off = 512; off > 0; off >>= 1){
		if(tid < off){ part[tid] += part[tid + off]; negs[tid] += negs[tid + off]; }
		__syncthreads();
	}
	if(tid == 0){
		scal[SCAL_LOGDET] = part[0];
		scal[SCAL_SIGN] = (negs[0] & 1) ? -1.0 : 1.0;
	}
}

__global__ void gj_copy_kernel(const double *__restrict__ W, double *__restrict__ out, size_t n, const double *__restrict__ scal)
{
	if(scal[SCAL_GJ_FLAG] != 0) return;   // the fallback inverts the untouched input in place
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) out[e] = W[e];
}

// Fallback: in-place Gauss-Jordan with partial pivoting on M (one workgroup).  Runs only when
// the fast path saw a multiplier above 1.
__global__ __launch_bounds__(1024) void gj_fallback_kernel(double *__restrict__ M, unsigned n, double *__restrict__ scal)
{
	extern __shared__ double fsh[];
	double *colg = fsh;                                   // n: multipliers of the step
	int *piv = reinterpret_cast<int *>(fsh + n);          // n
	__shared__ double rv[16];
	__shared__ int ri[16];
	__shared__ double ld_acc;
	__shared__ int neg_acc;
	if(scal[SCAL_GJ_FLAG] == 0) return;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	if(tid == 0){ ld_acc = 0; neg_acc = 0; }
	for(unsigned g = 0; g < n; g++){
		double best = -1;
		int bidx = 0x7fffffff;
		for(unsigned r = g + tid; r < n; r += 1024){
			const double a = fabs(M[(size_t)g * n + r]);
			if(a > best){ best = a; bidx = (int)r; }
		}
		for(int off = 32; off >= 1; off >>= 1){
			const double ob = __shfl_xor(best, off, 64);
			const int oi = __shfl_xor(bidx, off, 64);
			if(ob > best || (ob == best && oi < bidx)){ best = ob; bidx = oi; }
		}
		if(lane == 0){ rv[wave] = best; ri[wave] = bidx; }
		__syncthreads();
		best = rv[0]; bidx = ri[0];
		for(int w = 1; w < 16; w++) if(rv[w] > best || (rv[w] == best && ri[w] < bidx)){ best = rv[w]; bidx = ri[w]; }
		const unsigned p = bidx < (int)n ? (unsigned)bidx : g;
		if(tid == 0) piv[g] = (int)p;
		if(p != g)
			for(unsigned cc = tid; cc < n; cc += 1024){
				const double t = M[(size_t)cc * n + g];
				M[(size_t)cc * n + g] = M[(size_t)cc * n + p];
				M[(size_t)cc * n + p] = t;
			}
		__syncthreads();
		const double d = M[(size_t)g * n + g];
		for(unsigned r = tid; r < n; r += 1024) colg[r] = M[(size_t)g * n + r];
		if(tid == 0){ ld_acc += log(fabs(d)); neg_acc += (d < 0) + (p != g); }
		__syncthreads();
		const double inv = 1.0 / d;
		for(unsigned cc = tid; cc < n; cc += 1024) if(cc != g) M[(size_t)cc * n + g] *= inv;
		__syncthreads();
		for(unsigned cc = 0; cc < n; cc++){
			const double u = cc == g ? 0.0 : M[(size_t)cc * n + g];
			for(unsigned r = tid; r < n; r += 1024){
				if(r == g) continue;
				if(cc == g) M[(size_t)cc * n + r] = -colg[r] * inv;
				else M[(size_t)cc * n + r] -= colg[r] * u;
			}
		}
		if(tid == 0) M[(size_t)g * n + g] = inv;
		__syncthreads();
	}
	// undo the row exchanges on the columns, in reverse order
	for(int g = (int)n - 1; g >= 0; g--){
		const int p = piv[g];
		if(p != g)
			for(unsigned r = tid; r < n; r += 1024){
				const double t = M[(size_t)g * n + r];
				M[(size_t)g * n + r] = M[(size_t)p * n + r];
				M[(size_t)p * n + r] = t;
			}
		__syncthreads();
	}
	if(tid == 0){
		scal[SCAL_LOGDET] = ld_acc;
		scal[SCAL_SIGN] = (neg_acc & 1) ? -1.0 : 1.0;
		scal[SCAL_GJ_FALLBACKS] += 1;
	}
}

// M (column-major, N x N, FP64) <- M^{-1}; log|det| and sign into the device scalars.
int gj_logdet_inverse(ephmc_ctx *c, double *M)
{
	const unsigned n = c->ns;
	if(n > 4096){ set_error("gj: N = %u exceeds 4096", n); return EPHMC_ERR_UNSUPPORTED; }
	gj_clear_kernel<<<1, 1, 0, c->stream>>>(c->d_scal);
	EPHMC_LAUNCHED();
	const double *cur = M;
	double *buf[2] = {c->d_gj[0], c->d_gj[1]};
	int which = 0;
	const dim3 ugrid((n + 31) / 32, (n + 31) / 32);
	const bool fused = !(getenv("EPHMC_GJ_SPLIT") && atoi(getenv("EPHMC_GJ_SPLIT")) != 0);
	const dim3 sgrid((n + GJ_RB - 1) / GJ_RB, (n + GJ_CB - 1) / GJ_CB);
	for(unsigned k0 = 0; k0 < n; k0 += GJ_NB){
		const unsigned nbe = std::min<unsigned>(GJ_NB, n - k0);
		double *next = buf[which];
		if(fused){
			gj_step_kernel<<<sgrid, 256, 0, c->stream>>>(cur, next, n, k0, nbe, c->d_diag, c->d_piv, c->d_scal);
			EPHMC_LAUNCHED();
		}else{
			gj_panel_np_kernel<GJ_NB><<<(n + 255) / 256, 256, 0, c->stream>>>(cur, next, n, k0, nbe, c->d_diag, c->d_piv,
				c->d_scal);
			EPHMC_LAUNCHED();
			gj_update_kernel<GJ_NB><<<ugrid, 256, 0, c->stream>>>(cur, next, n, k0, nbe);
			EPHMC_LAUNCHED();
		}
		cur = next;
		which ^= 1;
	}
	gj_logdet_kernel<<<1, 1024, 0, c->stream>>>(c->d_diag, n, c->d_scal);
	EPHMC_LAUNCHED();
	gj_copy_kernel<<<1024, 256, 0, c->stream>>>(cur, M, c->mat, c->d_scal);
	EPHMC_LAUNCHED();
	const size_t lds = (size_t)n * (sizeof(double) + sizeof(int));
	(void)hipFuncSetAttribute((const void *)gj_fallback_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	gj_fallback_kernel<<<1, 1024, lds, c->stream>>>(M, n, c->d_scal);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

} // namespace ephmc
