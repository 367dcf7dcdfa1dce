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
//   fast path: blocked Gauss-Jordan without row exchanges, 32 columns per block K = [k0, k0+nb),
//     one launch per block over 64 x 64 output tiles.  Every workgroup inverts the diagonal
//     block D = A[K,K] (four waves, two pivots per barrier), forms the transform rows
//     E[i,:] = -A[i,K] D^{-1} (i not in K) or D^{-1} rows (i in K) of its rows, and writes its
//     tile of A' = [i not in K] A + E A[K,:] (columns inside K: E itself).  The LU multipliers
//     of D and of the rows below K are checked |l| <= 1
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

// 1/x: hardware reciprocal refined by two Newton steps
__device__ inline double rcp_nr(double x)
{
	double r = __builtin_amdgcn_rcp(x);
	double t = fma(-x, r, 1.0);
	r = fma(r, t, r);
	t = fma(-x, r, 1.0);
	return fma(r, t, r);
}

// Diagonal-block inversion.  D^{-1} by Gauss-Jordan on [D | 1] (32 x 64) without exchanges,
// two pivots (g, g+1) per step eliminated together through their 2 x 2 block
// P = [[a_gg, a_g,g+1], [a_g+1,g, a_g+1,g+1]]: pivot rows <- P^{-1} pivot rows, every other row
// r -= [a_rg a_r,g+1] (new pivot rows).  Wave w holds rows 8w..8w+7 in a[], lane c column c
// (lanes 32..63 end with D^{-1}); one barrier per step publishes the two pivot rows and the two
// pivot columns.  The pivot columns of every step are kept, and the LU multipliers of D (and the
// pivots U_gg) are formed from them afterwards: l[r][g] = a_rg / a_gg,
// U_hh = a_hh - l[h][g] a_gh, l[r][h] = (a_rh - l[r][g] a_gh) / U_hh.
// Writes dinv_out[r][c] = D^{-1}[r][c] and lm_out[m][j] = L[m][j] (unit lower triangular, zero
// elsewhere, NB x NB row-major), the pivots of rows kb.. to diag/piv, and raises the flag when a
// multiplier exceeds 1.  Called by all 256 threads of a workgroup.
struct GjInvLds {
	__align__(16) double prow[2][2][2 * GJ_NB];     // pivot rows of the step, by step parity
	__align__(16) double pcs[GJ_NB / 2][2][GJ_NB];  // pivot columns of every step
};

__device__ __forceinline__ void gj_invert_block(double (&a)[8], unsigned nbe, GjInvLds &sh, double *__restrict__ dinv_out,
		double *__restrict__ lm_out, double *__restrict__ diag, int *__restrict__ piv, unsigned kb, double *__restrict__ scal)
{
	constexpr int NB = GJ_NB;
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	for(int g = 0; g < (int)nbe; g += 2){
		const int h = g + 1, b = (g >> 1) & 1, gq = g & 7;
		if(wave == (g >> 3)){
			// rows gq, gq + 1 of this wave (selected without indexing a[] at run time)
#pragma unroll
			for(int q = 0; q < 8; q += 2)
				if(q == gq){
					sh.prow[b][0][lane] = a[q];
					sh.prow[b][1][lane] = a[q + 1];
				}
		}
		if(lane == g || lane == h){
			double *dst = &sh.pcs[g >> 1][lane - g][8 * wave];
#pragma unroll
			for(int q = 0; q < 8; q += 2) *(double2 *)(dst + q) = make_double2(a[q], a[q + 1]);
		}
		__syncthreads();
		const double *pg = sh.pcs[g >> 1][0], *ph = sh.pcs[g >> 1][1];
		const double pgg = pg[g], phg = pg[h], pgh = ph[g], phh = ph[h];
		const double idet = rcp_nr(fmac(pgg, phh, -pgh * phg));
		const double u = sh.prow[b][0][lane], v = sh.prow[b][1][lane];
		const double u2 = fmac(phh, u, -pgh * v) * idet, v2 = fmac(pgg, v, -phg * u) * idet;
		double cg[8], ch[8];
#pragma unroll
		for(int q = 0; q < 8; q += 2){
			const double2 x = *(const double2 *)(pg + 8 * wave + q), y = *(const double2 *)(ph + 8 * wave + q);
			cg[q] = x.x;
			cg[q + 1] = x.y;
			ch[q] = y.x;
			ch[q + 1] = y.y;
		}
#pragma unroll
		for(int q = 0; q < 8; q++){
			const int r = 8 * wave + q;
			a[q] = r == g ? u2 : (r == h ? v2 : fmac(-ch[q], v2, fmac(-cg[q], u2, a[q])));
		}
	}
	if(lane >= NB){
#pragma unroll
		for(int q = 0; q < 8; q++) dinv_out[(8 * wave + q) * NB + lane - NB] = a[q];
	}
	// L, entry (r, j) by thread tid + 256 i
	bool bad = false;
#pragma unroll
	for(int i = 0; i < NB * NB / 256; i++){
		const int e = tid + 256 * i, r = e / NB, j = e % NB;
		double l = r == j ? 1.0 : 0.0;
		if(r > j && r < (int)nbe){
			const int st = j >> 1, g = 2 * st;
			const double *pg = sh.pcs[st][0], *ph = sh.pcs[st][1];
			const double pgg = pg[g], lg = pg[r] / pgg;
			if(j == g) l = lg;
			else{
				const double uhh = ph[j] - (pg[j] / pgg) * ph[g];
				l = (ph[r] - lg * ph[g]) / uhh;
			}
			bad |= !(fabs(l) <= 1.0);
		}
		lm_out[e] = l;
	}
	if(tid < NB / 2 && 2 * tid < (int)nbe){
		const int g = 2 * tid, h = g + 1;
		const double *pg = sh.pcs[tid][0], *ph = sh.pcs[tid][1];
		diag[kb + g] = pg[g];
		piv[kb + g] = (int)(kb + g);
		if(h < (int)nbe){
			diag[kb + h] = ph[h] - (pg[h] / pg[g]) * ph[g];
			piv[kb + h] = (int)(kb + h);
		}
	}
	if(bad) scal[SCAL_GJ_FLAG] = 1;
}

// Inversion of the first diagonal block A[0:nbe, 0:nbe] (one workgroup).
__global__ __launch_bounds__(256) void gj_first_kernel(const double *__restrict__ A, unsigned n, unsigned nbe,
		double *__restrict__ dinv_out, double *__restrict__ lm_out, double *__restrict__ diag, int *__restrict__ piv,
		double *__restrict__ scal)
{
	constexpr int NB = GJ_NB;
	__shared__ GjInvLds sh;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	double a[8];
#pragma unroll
	for(int q = 0; q < 8; q++){
		const int r = 8 * wave + q;
		if(lane < NB) a[q] = (r < (int)nbe && lane < (int)nbe) ? A[(size_t)lane * n + r] : (r == lane ? 1.0 : 0.0);
		else a[q] = r == lane - NB ? 1.0 : 0.0;
	}
	gj_invert_block(a, nbe, sh, dinv_out, lm_out, diag, piv, 0, scal);
}

// Block step for K = [k0, k0+nbe), grid (row tiles + 1, column tiles) of 64 x 64 tiles, with
// D^{-1} and L of D = A[K,K] read from dinv_cur / lm_cur.
//
// Tile workgroups: load their operands (A[K, column block], A[row block, K], the tile), form the
// transform rows E[i,:] = -A[i,K] D^{-1} (rows inside K: the rows of D^{-1}) as 4 x 2 register
// blocks into a transposed LDS panel Et[m][row], and update 4 rows (16 apart) x 4 columns each:
// A'[rows, cols] = [rows not in K] A + E A[K, cols] (columns inside K: E itself).  Workgroups of
// the first column block check the LU multipliers of their rows below K (l = -E L, |l| <= 1).
//
// Lookahead workgroup (blockIdx = (row tiles, 0)): loads only the 32 x 32 operands of the next
// diagonal block D' = A'[K', K'], K' = [k0+nbe, ...), forms it with the same operations in the
// same order as the tile that holds it (so both hold the same bits), inverts it and writes
// D'^{-1}, L' and its pivots for the next step, so that no step waits for an inversion of its own.
constexpr int GJ_RB = 64, GJ_CB = 64, GJ_LD = 68;

__global__ __launch_bounds__(256) void gj_step_kernel(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe, const double *__restrict__ dinv_cur, const double *__restrict__ lm_cur,
		double *__restrict__ dinv_next, double *__restrict__ lm_next, double *__restrict__ diag, int *__restrict__ piv,
		double *__restrict__ scal)
{
	constexpr int NB = GJ_NB;
	__shared__ __align__(16) double dinv[NB][NB + 2];
	__shared__ double lm[NB][NB + 1];
	__shared__ __align__(16) double Art[NB][GJ_LD];         // A[row block, K], transposed
	__shared__ __align__(16) double Et[NB][GJ_LD];          // transform rows, transposed
	__shared__ __align__(16) double Ys[NB][GJ_LD];          // A[K, column block]
	__shared__ GjInvLds sh;
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const unsigned nrb = (n + GJ_RB - 1) / GJ_RB;
	const unsigned k1 = k0 + nbe;
	if(blockIdx.x == nrb){
		// ---- lookahead: D' = A[K',K'] + E[K',:] A[K,K'] with E[K',:] = -A[K',K] D^{-1}, entry by
		// entry the same fma sequence as the tile that holds it (m ascending over all NB columns)
		if(blockIdx.y != 0 || k1 >= n) return;
		const unsigned nb1 = min((unsigned)NB, n - k1);
		// every global load is issued before the first LDS store, so they share one round trip
		double av[NB * NB / 256], dv[NB * NB / 256], xv[NB * NB / 256], yv[NB * NB / 256];
#pragma unroll
		for(int q = 0; q < NB * NB / 256; q++){
			const int e = tid + 256 * q, i = e % NB, j = e / NB;     // i: row (fastest), j: column / m
			dv[q] = dinv_cur[e];
			xv[q] = (j < (int)nbe && i < (int)nb1) ? A[(size_t)(k0 + j) * n + k1 + i] : 0.0;
			yv[q] = (i < (int)nbe && j < (int)nb1) ? A[(size_t)(k1 + j) * n + k0 + i] : 0.0;
			av[q] = (i < (int)nb1 && j < (int)nb1) ? A[(size_t)(k1 + j) * n + k1 + i] : 0.0;
		}
#pragma unroll
		for(int q = 0; q < NB * NB / 256; q++){
			const int e = tid + 256 * q, i = e % NB, j = e / NB;
			dinv[j][i] = dv[q];
			Art[j][i] = xv[q];
			Ys[i][j] = yv[q];
		}
		__syncthreads();
#pragma unroll
		for(int q = 0; q < NB * NB / 256; q++){
			const int e = tid + 256 * q, i = e % NB, j = e / NB;
			double x = 0;
#pragma unroll 8
			for(int m = 0; m < NB; m++) x = fmac(Art[m][i], dinv[m][j], x);
			Et[j][i] = -x;
		}
		__syncthreads();
		double (*Dn)[GJ_LD] = Art;           // free once E' is formed
#pragma unroll
		for(int q = 0; q < NB * NB / 256; q++){
			const int e = tid + 256 * q, i = e % NB, j = e / NB;
			double x = av[q];
#pragma unroll 8
			for(int m = 0; m < NB; m++) x = fmac(Et[m][i], Ys[m][j], x);
			Dn[i][j] = x;
		}
		__syncthreads();
		double a[8];
#pragma unroll
		for(int q = 0; q < 8; q++){
			const int r = 8 * wave + q;
			if(lane < NB) a[q] = (r < (int)nb1 && lane < (int)nb1) ? Dn[r][lane] : (r == lane ? 1.0 : 0.0);
			else a[q] = r == lane - NB ? 1.0 : 0.0;
		}
		gj_invert_block(a, nb1, sh, dinv_next, lm_next, diag, piv, k1, scal);
		return;
	}
	const unsigned r0 = blockIdx.x * GJ_RB, c0 = blockIdx.y * GJ_CB;
	// ---- loads: D^{-1}, L (first column block), panels and tile values, all issued before the
	// first LDS store so that they share one round trip
	const bool first = blockIdx.y == 0;
	double dv[NB * NB / 256], lv[NB * NB / 256], ysv[NB * GJ_CB / 256], arv[GJ_RB * NB / 256];
#pragma unroll
	for(int i = 0; i < NB * NB / 256; i++){
		const int e = tid + 256 * i;
		dv[i] = dinv_cur[e];
		lv[i] = first ? lm_cur[e] : 0.0;
	}
#pragma unroll
	for(int j = 0; j < NB * GJ_CB / 256; j++){
		const int e = tid + 256 * j, m = e % NB, cc = e / NB;
		const unsigned col = c0 + cc;
		ysv[j] = (m < (int)nbe && col < n) ? A[(size_t)col * n + k0 + m] : 0.0;
	}
#pragma unroll
	for(int j = 0; j < GJ_RB * NB / 256; j++){
		const int e = tid + 256 * j, rr = e % GJ_RB, m = e / GJ_RB;
		const unsigned i = r0 + rr;
		arv[j] = (m < (int)nbe && i < n) ? A[(size_t)(k0 + m) * n + i] : 0.0;
	}
	// this thread's tile block: rows tr + 16 i, columns 4 tc + j (16 lanes read 16 consecutive rows)
	const int tr = tid & 15, tc = tid >> 4;
	double acc[4][4];
#pragma unroll
	for(int i = 0; i < 4; i++){
		const unsigned row = r0 + tr + 16 * i;
		const bool rowK = row >= k0 && row < k1;
#pragma unroll
		for(int j = 0; j < 4; j++){
			const unsigned col = c0 + 4 * tc + j;
			acc[i][j] = (row < n && col < n && !rowK) ? A[(size_t)col * n + row] : 0.0;
		}
	}
#pragma unroll
	for(int i = 0; i < NB * NB / 256; i++){
		const int e = tid + 256 * i;
		dinv[e / NB][e % NB] = dv[i];
		if(first) lm[e / NB][e % NB] = lv[i];
	}
#pragma unroll
	for(int j = 0; j < NB * GJ_CB / 256; j++){
		const int e = tid + 256 * j;
		Ys[e % NB][e / NB] = ysv[j];
	}
#pragma unroll
	for(int j = 0; j < GJ_RB * NB / 256; j++){
		const int e = tid + 256 * j;
		Art[e / GJ_RB][e % GJ_RB] = arv[j];
	}
	__syncthreads();
	// ---- transform rows: thread (rows 4 rq + i, columns 2 cq + jj)
	{
		const int rq = tid & 15, cq = tid >> 4;
		double e[4][2] = {};
#pragma unroll 8
		for(int m = 0; m < NB; m++){
			const double2 x0 = *(const double2 *)&Art[m][4 * rq], x1 = *(const double2 *)&Art[m][4 * rq + 2];
			const double2 dv = *(const double2 *)&dinv[m][2 * cq];
			const double xr[4] = {x0.x, x0.y, x1.x, x1.y};
#pragma unroll
			for(int i = 0; i < 4; i++){
				e[i][0] = fmac(xr[i], dv.x, e[i][0]);
				e[i][1] = fmac(xr[i], dv.y, e[i][1]);
			}
		}
#pragma unroll
		for(int i = 0; i < 4; i++){
			const unsigned row = r0 + 4 * rq + i;
			const bool inK = row >= k0 && row < k1;
#pragma unroll
			for(int jj = 0; jj < 2; jj++)
				Et[2 * cq + jj][4 * rq + i] = inK ? dinv[row - k0][2 * cq + jj] : -e[i][jj];
		}
	}
	__syncthreads();
	// LU multipliers of the rows below K (first column block only): thread (row, columns
	// jg + 4 t), y_j = sum_m E[row][m] L[m][j]
	if(blockIdx.y == 0){
		const int row = tid & 63, jg = tid >> 6;
		const unsigned i = r0 + row;
		if(i < n && i >= k1){
			double er[NB];
#pragma unroll
			for(int m = 0; m < NB; m++) er[m] = Et[m][row];
			bool bad = false;
#pragma unroll
			for(int t = 0; t < NB / 4; t++){
				const int j = jg + 4 * t;
				double y = 0;
#pragma unroll
				for(int m = 0; m < NB; m++) y = fmac(er[m], lm[m][j], y);
				bad |= j < (int)nbe && !(fabs(y) <= 1.0);
			}
			if(bad) scal[SCAL_GJ_FLAG] = 1;
		}
	}
	// ---- tile update: acc += E[rows, :] A[K, cols]
#pragma unroll 8
	for(int m = 0; m < NB; m++){
		const double2 y0 = *(const double2 *)&Ys[m][4 * tc], y1 = *(const double2 *)&Ys[m][4 * tc + 2];
		const double ev[4] = {Et[m][tr], Et[m][tr + 16], Et[m][tr + 32], Et[m][tr + 48]};
		const double yv[4] = {y0.x, y0.y, y1.x, y1.y};
#pragma unroll
		for(int i = 0; i < 4; i++)
#pragma unroll
			for(int j = 0; j < 4; j++) acc[i][j] = fmac(ev[i], yv[j], acc[i][j]);
	}
#pragma unroll
	for(int j = 0; j < 4; j++){
		const unsigned col = c0 + 4 * tc + j;
		if(col >= n) continue;
		const bool colK = col >= k0 && col < k1;
#pragma unroll
		for(int i = 0; i < 4; i++){
			const unsigned row = r0 + tr + 16 * i;
			if(row < n) An[(size_t)col * n + row] = colK ? Et[col - k0][tr + 16 * i] : acc[i][j];
		}
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
	// D^{-1} and L of the current and the next diagonal block (double-buffered by step parity)
	double *dinv_g[2] = {c->d_dlu, c->d_dlu + GJ_NB * GJ_NB}, *lm_g[2] = {c->d_dlu + 2 * GJ_NB * GJ_NB, c->d_dlu + 3 * GJ_NB * GJ_NB};
	gj_first_kernel<<<1, 256, 0, c->stream>>>(M, n, std::min<unsigned>(GJ_NB, n), dinv_g[0], lm_g[0], c->d_diag, c->d_piv, c->d_scal);
	EPHMC_LAUNCHED();
	const dim3 sgrid((n + GJ_RB - 1) / GJ_RB + 1, (n + GJ_CB - 1) / GJ_CB);
	int par = 0;
	for(unsigned k0 = 0; k0 < n; k0 += GJ_NB){
		const unsigned nbe = std::min<unsigned>(GJ_NB, n - k0);
		double *next = buf[which];
		gj_step_kernel<<<sgrid, 256, 0, c->stream>>>(cur, next, n, k0, nbe, dinv_g[par], lm_g[par], dinv_g[par ^ 1], lm_g[par ^ 1],
			c->d_diag, c->d_piv, c->d_scal);
		EPHMC_LAUNCHED();
		cur = next;
		which ^= 1;
		par ^= 1;
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
