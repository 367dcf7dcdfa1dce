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
//     block D = A[K,K] (four waves, four pivots per barrier), forms the transform rows
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
// four pivots g..g+3 per step eliminated together through their 4 x 4 block P (rows and
// columns g..g+3): with P = L_P U_P (no exchanges), every column's new pivot rows are
// nr = P^{-1} u (u: the column's entries in the pivot rows) by forward and back substitution,
// and every other row r -= a_r,[g..g+3] nr.  Wave w holds rows 8w..8w+7 in a[], lane c column c
// (lanes 32..63 end with D^{-1}); one barrier per step publishes the four pivot rows, from which
// every thread forms P and its factorisation (the same values everywhere, so no second barrier is
// needed to share it).  A row's entries in the pivot columns (a_r,[g..g+3]) are read from the
// lanes g..g+3 of its own wave.  Each lane c < 32 keeps its column as it stood at the step that
// pivots on it; the kept columns go to LDS after the loop.  After the loop the pivots U_gg and the LU multipliers of
// D are formed from the kept pivot columns: row r's multipliers of step s are
// l[r][g+i] = (a_r,g+i - sum_{q<i} l[r][g+q] U_P[q][i]) / U_P[i][i]  (they only feed the
// |l| <= 1 check).  Writes dinv_out[r][c] = D^{-1}[r][c] and lm_out[m][j] = L[m][j] (unit lower
// triangular, zero elsewhere, NB x NB row-major), the pivots of rows kb.. to diag/piv, and raises
// the flag when a multiplier exceeds 1.  Called by all 256 threads of a workgroup.
struct GjInvLds {
	__align__(16) double prow[2][4][2 * GJ_NB];     // pivot rows of the step, by step parity
	__align__(16) double pcs[GJ_NB / 4][4][GJ_NB + 2];  // pivot columns of every step (padded: the
	                                                // multiplier pass reads 8 steps at once)
};

// L and U of the 4 x 4 pivot block P (P[m][i] = row g+m, column g+i) without exchanges:
// lo[m][k] (m > k) the multipliers, up[k][i] (i > k) U, rp[k] = 1 / U[k][k], pv[k] = U[k][k]
__device__ __forceinline__ void gj_lu4_of(double (&M)[4][4], double (&lo)[4][4], double (&up)[4][4], double (&rp)[4],
		double (&pv)[4])
{
#pragma unroll
	for(int k = 0; k < 4; k++){
		pv[k] = M[k][k];
		rp[k] = rcp_nr(M[k][k]);
#pragma unroll
		for(int i = k + 1; i < 4; i++) up[k][i] = M[k][i];
#pragma unroll
		for(int m = k + 1; m < 4; m++){
			const double l = M[m][k] * rp[k];
			lo[m][k] = l;
#pragma unroll
			for(int i = k + 1; i < 4; i++) M[m][i] = fmac(-l, M[k][i], M[m][i]);
		}
	}
}

// P of step s from the kept pivot columns
__device__ __forceinline__ void gj_lu4(const GjInvLds &sh, int s, double (&lo)[4][4], double (&up)[4][4], double (&rp)[4],
		double (&pv)[4])
{
	const int g = 4 * s;
	double M[4][4];
#pragma unroll
	for(int i = 0; i < 4; i++){
		const double2 x0 = *(const double2 *)&sh.pcs[s][i][g], x1 = *(const double2 *)&sh.pcs[s][i][g + 2];
		M[0][i] = x0.x;
		M[1][i] = x0.y;
		M[2][i] = x1.x;
		M[3][i] = x1.y;
	}
	gj_lu4_of(M, lo, up, rp, pv);
}

// P of the current step from its published pivot rows (the same values as its pivot columns)
__device__ __forceinline__ void gj_lu4_rows(const GjInvLds &sh, int b, int g, double (&lo)[4][4], double (&up)[4][4],
		double (&rp)[4], double (&pv)[4])
{
	double M[4][4];
#pragma unroll
	for(int m = 0; m < 4; m++){
		const double2 x0 = *(const double2 *)&sh.prow[b][m][g], x1 = *(const double2 *)&sh.prow[b][m][g + 2];
		M[m][0] = x0.x;
		M[m][1] = x0.y;
		M[m][2] = x1.x;
		M[m][3] = x1.y;
	}
	gj_lu4_of(M, lo, up, rp, pv);
}

// the value of v in lane `src` of this wave (src uniform across the wave)
__device__ __forceinline__ double lane_value(double v, int src)
{
	const long long b = __double_as_longlong(v);
	const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), src);
	const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
	return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ void gj_invert_block(double (&a)[8], unsigned nbe, GjInvLds &sh, double *__restrict__ dinv_out,
		double *__restrict__ lm_out, double *__restrict__ diag, int *__restrict__ piv, unsigned kb, double *__restrict__ scal)
{
	constexpr int NB = GJ_NB;
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	double kept[8];
#pragma unroll
	for(int q = 0; q < 8; q++) kept[q] = 0;
	for(int s = 0; s < NB / 4; s++){
		const int g = 4 * s, b = s & 1, gq = g & 7;
		if(g >= (int)nbe) break;
		if(wave == (g >> 3)){
			// rows gq .. gq + 3 of this wave (selected without indexing a[] at run time)
#pragma unroll
			for(int q = 0; q < 8; q += 4)
				if(q == gq){
#pragma unroll
					for(int i = 0; i < 4; i++) sh.prow[b][i][lane] = a[q + i];
				}
		}
		if((lane >> 2) == s){
#pragma unroll
			for(int q = 0; q < 8; q++) kept[q] = a[q];
		}
		__syncthreads();
		double lo[4][4], up[4][4], rp[4], pv[4];
		gj_lu4_rows(sh, b, g, lo, up, rp, pv);
		// (no store of U_P or the pivots here: a store inside the loop delays the next barrier; they
		// are formed again from the kept pivot columns after the loop)
		// new pivot rows of this column: y = L^{-1} u, nr = U^{-1} y
		double u[4];
#pragma unroll
		for(int i = 0; i < 4; i++) u[i] = sh.prow[b][i][lane];
		double y[4], nr[4];
#pragma unroll
		for(int m = 0; m < 4; m++){
			double x = u[m];
#pragma unroll
			for(int k = 0; k < m; k++) x = fmac(-lo[m][k], y[k], x);
			y[m] = x;
		}
#pragma unroll
		for(int m = 3; m >= 0; m--){
			double x = y[m];
#pragma unroll
			for(int i = m + 1; i < 4; i++) x = fmac(-up[m][i], nr[i], x);
			nr[m] = x * rp[m];
		}
		double c[8][4];
#pragma unroll
		for(int i = 0; i < 4; i++)
#pragma unroll
			for(int q = 0; q < 8; q++) c[q][i] = lane_value(a[q], g + i);
#pragma unroll
		for(int q = 0; q < 8; q++){
			const int r = 8 * wave + q;
			double x = a[q];
#pragma unroll
			for(int i = 0; i < 4; i++) x = fmac(-c[q][i], nr[i], x);
			a[q] = r == g ? nr[0] : (r == g + 1 ? nr[1] : (r == g + 2 ? nr[2] : (r == g + 3 ? nr[3] : x)));
		}
	}
	if(lane >= NB){
#pragma unroll
		for(int q = 0; q < 8; q++) dinv_out[(8 * wave + q) * NB + lane - NB] = a[q];
	}
	if(lane < (int)nbe){
		double *dst = &sh.pcs[lane >> 2][lane & 3][8 * wave];
#pragma unroll
		for(int q = 0; q < 8; q += 2) *(double2 *)(dst + q) = make_double2(kept[q], kept[q + 1]);
	}
	__syncthreads();
	// the pivots and the multipliers: thread column j refactors the pivot block of its step s
	// (the same arithmetic as inside the loop, so the same values) and takes row j & 3 of U_P
	bool bad = false;
	{
		const int j = tid % NB, s = j >> 2, ii = j & 3;
		double my_u[4][4], rp[4], pv[4];
		{
			double lo[4][4], up[4][4];
			gj_lu4(sh, s, lo, up, rp, pv);
#pragma unroll
			for(int q = 0; q < 4; q++)
#pragma unroll
				for(int i = 0; i < 4; i++) my_u[q][i] = i == q ? rp[q] : (i > q ? up[q][i] : 0.0);
		}
		if(tid < (int)nbe){
			diag[kb + tid] = ii == 0 ? pv[0] : (ii == 1 ? pv[1] : (ii == 2 ? pv[2] : pv[3]));
			piv[kb + tid] = (int)(kb + tid);
		}
#pragma unroll
		for(int i = 0; i < NB * NB / 256; i++){
			const int e = tid + 256 * i, r = e / NB;
			double l = r == j ? 1.0 : 0.0;
			if(r > j && r < (int)nbe){
				// l[r][g+q] = (a_r,g+q - sum_{p<q} l[r][g+p] U[p][q]) / U[q][q]
				double lr[4];
#pragma unroll
				for(int q = 0; q < 4; q++){
					if(q > ii) break;
					double x = sh.pcs[s][q][r];
#pragma unroll
					for(int p = 0; p < q; p++) x = fmac(-lr[p], my_u[p][q], x);
					lr[q] = x * my_u[q][q];
				}
				l = ii == 0 ? lr[0] : (ii == 1 ? lr[1] : (ii == 2 ? lr[2] : lr[3]));
				bad |= !(fabs(l) <= 1.0);
			}
			lm_out[e] = l;
		}
	}
	if(bad) scal[SCAL_GJ_FLAG] = 1;
}

// Inversion of the first diagonal block A[0:nbe, 0:nbe] (one workgroup).
__device__ __forceinline__ void gj_first_body(const double *__restrict__ A, unsigned n, unsigned nbe,
		double *__restrict__ dinv_out, double *__restrict__ lm_out, double *__restrict__ diag, int *__restrict__ piv,
		double *__restrict__ scal)
{
	constexpr int NB = GJ_NB;
	__shared__ GjInvLds sh;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	if(threadIdx.x == 0){
		scal[SCAL_GJ_FLAG] = 0;     // no multiplier above 1 seen yet
		scal[SCAL_SINGULAR] = 0;
		scal[SCAL_SCHUR_FAIL] = 0;
	}
	__syncthreads();
	double a[8];
#pragma unroll
	for(int q = 0; q < 8; q++){
		const int r = 8 * wave + q;
		if(lane < NB) a[q] = (r < (int)nbe && lane < (int)nbe) ? A[(size_t)lane * n + r] : (r == lane ? 1.0 : 0.0);
		else a[q] = r == lane - NB ? 1.0 : 0.0;
	}
	gj_invert_block(a, nbe, sh, dinv_out, lm_out, diag, piv, 0, scal);
}

__global__ __launch_bounds__(256) void gj_first_kernel(const double *__restrict__ A, unsigned n, unsigned nbe,
		double *__restrict__ dinv_out, double *__restrict__ lm_out, double *__restrict__ diag, int *__restrict__ piv,
		double *__restrict__ scal)
{
	gj_first_body(A, n, nbe, dinv_out, lm_out, diag, piv, scal);
}

// Block step for K = [k0, k0+nbe), grid (row tiles + 1, column tiles) of 64 x 64 tiles, with
// D^{-1} and L of D = A[K,K] read from dinv_cur / lm_cur.
//
// Tile workgroups: load their operands (A[K, column block], A[row block, K], the tile), form the
// transform rows E[i,:] = -A[i,K] D^{-1} (rows inside K: the rows of D^{-1}) as 4 x 2 register
// blocks into a transposed LDS panel Et[m][row], and update 4 rows (16 apart) x 4 columns each:
// A'[rows, cols] = [rows not in K] A + E A[K, cols] (columns inside K: E itself).  The tiles of
// a row block share the check of its LU multipliers below K (l = -E L, |l| <= 1): tile y checks
// the columns j = y, y + ncb, ... of L.
//
// Lookahead workgroup (blockIdx = (row tiles, 0)): loads only the 32 x 32 operands of the next
// diagonal block D' = A'[K', K'], K' = [k0+nbe, ...), forms it on the matrix cores (equal to
// the tile that holds it up to rounding), inverts it and writes D'^{-1}, L' and its pivots for
// the next step, so that no step waits for an inversion of its own.
constexpr int GJ_RB = 64, GJ_CB = 64, GJ_LD = 68;

__device__ __forceinline__ void gj_step_body(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe, const double *__restrict__ dinv_cur, const double *__restrict__ lm_cur,
		double *__restrict__ dinv_next, double *__restrict__ lm_next, double *__restrict__ diag, int *__restrict__ piv,
		double *__restrict__ scal, double *__restrict__ out_final, unsigned bx, unsigned by, unsigned ncb_grid)
{
	constexpr int NB = GJ_NB;
	__shared__ __align__(16) double dinv[NB][NB + 2];
	__shared__ double lm[NB][NB + 1];
	__shared__ __align__(16) double Art[NB][GJ_LD];         // A[row block, K], transposed
	__shared__ __align__(16) double Et[NB][GJ_LD];          // transform rows, transposed
	__shared__ __align__(16) double Ys[NB][GJ_LD];          // A[K, column block]
	// the lookahead's inversion scratch reuses Et (free once D' is formed), which keeps the
	// kernel at about 69 KB of LDS: two workgroups per CU, so the 257 workgroups of a 1024 x 1024
	// step all start at once
	static_assert(sizeof(GjInvLds) <= sizeof(double) * NB * GJ_LD, "inversion scratch exceeds Et");
	GjInvLds &sh = *reinterpret_cast<GjInvLds *>(&Et[0][0]);
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const unsigned nrb = (n + GJ_RB - 1) / GJ_RB;
	const unsigned k1 = k0 + nbe;
	if(bx == nrb){
		// ---- lookahead: D' = A[K',K'] + E[K',:] A[K,K'] with E[K',:] = -A[K',K] D^{-1}, both
		// products on the FP64 matrix cores (v_mfma_f64_16x16x4): wave w forms the 16 x 16 quadrant
		// rows 16 qi + m, columns 16 qj + n (qi = w / 2, qj = w % 2) in 8 steps of k = 4.  The tile
		// that holds A'[K',K'] sums in another order, so D' may differ from it in the last bits: the
		// inversion uses D', the next step's panels use the tile's values (rounding-level difference)
		if(by != 0) return;
		if(k1 >= n){
			// last step: no next block; with out_final this workgroup sums log|det| and its sign
			// from the pivots (the flag is final: the last block's multipliers were checked when it
			// was inverted)
			if(out_final && scal[SCAL_GJ_FLAG] == 0){
				double *part = &Art[0][0];
				int *negs = reinterpret_cast<int *>(&Ys[0][0]);
				double sl = 0;
				int neg = 0;
				for(unsigned i = tid; i < n; i += 256){
					const double d = diag[i];
					sl += log(fabs(d));
					neg += d < 0;
				}
				part[tid] = sl;
				negs[tid] = neg;
				__syncthreads();
				for(int off = 128; off > 0; off >>= 1){
					if(tid < off){ part[tid] += part[tid + off]; negs[tid] += negs[tid + off]; }
					__syncthreads();
				}
				if(tid == 0){
					scal[SCAL_LOGDET] = part[0];
					scal[SCAL_SIGN] = (negs[0] & 1) ? -1.0 : 1.0;
				}
			}
			return;
		}
		const unsigned nb1 = min((unsigned)NB, n - k1);
		typedef double gj_v4d __attribute__((ext_vector_type(4)));
		const int m16 = lane & 15, k4 = lane >> 4, qi = wave >> 1, qj = wave & 1;
		// every global load is issued before the first LDS store, so they share one round trip;
		// A[K',K'] is read straight into the accumulator layout (row 16 qi + 4 e + k4, column 16 qj + m16)
		double dv[NB * NB / 256], xv[NB * NB / 256], yv[NB * NB / 256], avm[4];
#pragma unroll
		for(int q = 0; q < NB * NB / 256; q++){
			const int e = tid + 256 * q, i = e % NB, j = e / NB;     // i: row (fastest), j: column / m
			dv[q] = dinv_cur[e];
			xv[q] = (j < (int)nbe && i < (int)nb1) ? A[(size_t)(k0 + j) * n + k1 + i] : 0.0;
			yv[q] = (i < (int)nbe && j < (int)nb1) ? A[(size_t)(k1 + j) * n + k0 + i] : 0.0;
		}
#pragma unroll
		for(int e = 0; e < 4; e++){
			const int i = 16 * qi + 4 * e + k4, j = 16 * qj + m16;
			avm[e] = (i < (int)nb1 && j < (int)nb1) ? A[(size_t)(k1 + j) * n + k1 + i] : 0.0;
		}
#pragma unroll
		for(int q = 0; q < NB * NB / 256; q++){
			const int e = tid + 256 * q, i = e % NB, j = e / NB;
			dinv[j][i] = dv[q];
			Art[j][i] = xv[q];
			Ys[i][j] = yv[q];
		}
		__syncthreads();
		gj_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
		for(int kb = 0; kb < NB / 4; kb++){
			const int k = 4 * kb + k4;
			acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Art[k][16 * qi + m16], dinv[k][16 * qj + m16], acc, 0, 0, 0);
		}
#pragma unroll
		for(int e = 0; e < 4; e++) Et[16 * qj + m16][16 * qi + 4 * e + k4] = -acc[e];
		__syncthreads();
		double (*Dn)[GJ_LD] = Art;           // free once E' is formed
		gj_v4d dac = {avm[0], avm[1], avm[2], avm[3]};
#pragma unroll
		for(int kb = 0; kb < NB / 4; kb++){
			const int k = 4 * kb + k4;
			dac = __builtin_amdgcn_mfma_f64_16x16x4f64(Et[k][16 * qi + m16], Ys[k][16 * qj + m16], dac, 0, 0, 0);
		}
#pragma unroll
		for(int e = 0; e < 4; e++) Dn[16 * qi + 4 * e + k4][16 * qj + m16] = dac[e];
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
	const unsigned r0 = bx * GJ_RB, c0 = by * GJ_CB;
	// the last step writes the inverse straight to out_final unless the pivoted fallback will run
	// (it inverts the untouched input there)
	double *__restrict__ dst = (out_final && scal[SCAL_GJ_FLAG] == 0) ? out_final : An;
	// ---- loads: D^{-1}, L (first column block), panels and tile values, all issued before the
	// first LDS store so that they share one round trip
	// the LU multipliers of a row block are checked by all its tiles: tile y takes the columns
	// j = y, y + ncb, ... of L (lm[m][jj] holds column y + ncb jj)
	const unsigned ncb = ncb_grid;
	double dv[NB * NB / 256], lv[NB * NB / 256], ysv[NB * GJ_CB / 256], arv[GJ_RB * NB / 256];
#pragma unroll
	for(int i = 0; i < NB * NB / 256; i++){
		const int e = tid + 256 * i, m = e % NB, jj = e / NB;
		const unsigned j = by + ncb * (unsigned)jj;
		dv[i] = dinv_cur[e];
		lv[i] = j < (unsigned)NB ? lm_cur[m * NB + j] : 0.0;
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
	// the tile on the FP64 matrix cores in transposed form, T^T[col][row] += sum_m Y[m][col] E[row][m]
	// (v_mfma_f64_16x16x4): wave w holds the 16 x 16 blocks of columns 16 (2 (w >> 1) + cb) and
	// rows 16 (2 (w & 1) + rb); lane l has rows 16 rb' + l % 16 (16 lanes read 16 consecutive
	// rows) and columns 16 cb' + 4 e + l / 16 in element e
	typedef double gj_v4d __attribute__((ext_vector_type(4)));
	const int m16 = lane & 15, k4 = lane >> 4, wcb = 2 * (wave >> 1), wrb = 2 * (wave & 1);
	gj_v4d acc[2][2];
#pragma unroll
	for(int cb = 0; cb < 2; cb++)
#pragma unroll
		for(int rb = 0; rb < 2; rb++){
			const unsigned row = r0 + 16 * (wrb + rb) + m16;
			const bool rowK = row >= k0 && row < k1;
#pragma unroll
			for(int e = 0; e < 4; e++){
				const unsigned col = c0 + 16 * (wcb + cb) + 4 * e + k4;
				acc[cb][rb][e] = (row < n && col < n && !rowK) ? A[(size_t)col * n + row] : 0.0;
			}
		}
#pragma unroll
	for(int i = 0; i < NB * NB / 256; i++){
		const int e = tid + 256 * i;
		dinv[e / NB][e % NB] = dv[i];
		lm[e % NB][e / NB] = lv[i];
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
	// ---- transform rows on the matrix cores, transposed: E^T[j][i] = sum_m D^{-1}[m][j] X[i][m]
	// (X[i][m] = Art[m][i]); wave w forms columns j of block w & 1 for the rows of blocks
	// 2 (w >> 1) + ib; rows inside K take the rows of D^{-1}
	{
		const int jb = wave & 1, ib0 = 2 * (wave >> 1);
		gj_v4d ea[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
		for(int kb = 0; kb < NB / 4; kb++){
			const int k = 4 * kb + k4;
			const double av = dinv[k][16 * jb + m16];
#pragma unroll
			for(int ib = 0; ib < 2; ib++)
				ea[ib] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Art[k][16 * (ib0 + ib) + m16], ea[ib], 0, 0, 0);
		}
#pragma unroll
		for(int ib = 0; ib < 2; ib++){
			const int il = 16 * (ib0 + ib) + m16;
			const unsigned row = r0 + il;
			const bool inK = row >= k0 && row < k1;
#pragma unroll
			for(int e = 0; e < 4; e++){
				const int j = 16 * jb + 4 * e + k4;
				Et[j][il] = inK ? dinv[row - k0][j] : -ea[ib][e];
			}
		}
	}
	__syncthreads();
	// LU multipliers of the rows below K: thread (row, columns y + ncb (jg + 4 t)),
	// y_j = sum_m E[row][m] L[m][j]
	{
		const int row = tid & 63, jg = tid >> 6;
		const unsigned i = r0 + row;
		if(i < n && i >= k1 && by + ncb * (unsigned)jg < (unsigned)NB){
			bool bad = false;
			for(int jj = jg; by + ncb * (unsigned)jj < (unsigned)NB; jj += 4){
				const unsigned j = by + ncb * (unsigned)jj;
				double y = 0;
#pragma unroll 8
				for(int m = 0; m < NB; m++) y = fmac(Et[m][row], lm[m][jj], y);
				bad |= j < nbe && !(fabs(y) <= 1.0);
			}
			if(bad) scal[SCAL_GJ_FLAG] = 1;
		}
	}
	// ---- tile update: T^T[col][row] += sum_m Y[m][col] E[row][m] (Ys[m][col], Et[m][row])
#pragma unroll
	for(int kb = 0; kb < NB / 4; kb++){
		const int k = 4 * kb + k4;
		double yv[2], ev[2];
#pragma unroll
		for(int q = 0; q < 2; q++){
			yv[q] = Ys[k][16 * (wcb + q) + m16];
			ev[q] = Et[k][16 * (wrb + q) + m16];
		}
#pragma unroll
		for(int cb = 0; cb < 2; cb++)
#pragma unroll
			for(int rb = 0; rb < 2; rb++)
				acc[cb][rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(yv[cb], ev[rb], acc[cb][rb], 0, 0, 0);
	}
#pragma unroll
	for(int cb = 0; cb < 2; cb++)
#pragma unroll
		for(int e = 0; e < 4; e++){
			const int cl = 16 * (wcb + cb) + 4 * e + k4;
			const unsigned col = c0 + cl;
			if(col >= n) continue;
			const bool colK = col >= k0 && col < k1;
#pragma unroll
			for(int rb = 0; rb < 2; rb++){
				const int rl = 16 * (wrb + rb) + m16;
				const unsigned row = r0 + rl;
				if(row < n) dst[(size_t)col * n + row] = colK ? Et[col - k0][rl] : acc[cb][rb][e];
			}
		}
}

__global__ __launch_bounds__(256) void gj_step_kernel(const double *__restrict__ A, double *__restrict__ An,
		unsigned n, unsigned k0, unsigned nbe, const double *__restrict__ dinv_cur, const double *__restrict__ lm_cur,
		double *__restrict__ dinv_next, double *__restrict__ lm_next, double *__restrict__ diag, int *__restrict__ piv,
		double *__restrict__ scal, double *__restrict__ out_final)
{
	gj_step_body(A, An, n, k0, nbe, dinv_cur, lm_cur, dinv_next, lm_next, diag, piv, scal, out_final, blockIdx.x, blockIdx.y,
		gridDim.y);
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
// the fast path saw a multiplier above 1 (or a non-finite one).  The row interchanges go to
// piv_out (LAPACK's sequential form, 0-based: row g was exchanged with row piv_out[g] at step g).
// A column without a nonzero candidate pivot makes the matrix singular: log|det| = -inf (the
// reference's sum of log|U_ii| with a zero U_ii, organic_dense_matrix.c:582-584), the step is
// recorded in SCAL_SINGULAR, and M is filled with NaN, so that a force or a Metropolis step built
// on it rejects instead of using a partial inverse.
__device__ __forceinline__ void gj_fallback_body(double *__restrict__ M, unsigned n, double *__restrict__ scal,
		int *__restrict__ piv_out)
{
	extern __shared__ double fsh[];
	double *colg = fsh;                                   // n: multipliers of the step
	double *rowg = fsh + n;                               // n: the scaled pivot row of the step
	int *piv = reinterpret_cast<int *>(fsh + 2 * n);      // n
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
		if(best == 0){
			// every candidate in column g is zero (best is the same in every thread): singular
			for(size_t e = tid; e < (size_t)n * n; e += 1024) M[e] = __longlong_as_double(0x7ff8000000000000LL);
			if(tid == 0){
				scal[SCAL_LOGDET] = -HUGE_VAL;
				scal[SCAL_SIGN] = 1.0;
				scal[SCAL_SINGULAR] = g + 1;
				scal[SCAL_GJ_FALLBACKS] += 1;
			}
			for(unsigned r = g + tid; r < n; r += 1024) piv_out[r] = (int)r;
			return;
		}
		const unsigned p = bidx < (int)n ? (unsigned)bidx : g;
		if(tid == 0){
			piv[g] = (int)p;
			piv_out[g] = (int)p;
		}
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
		for(unsigned cc = tid; cc < n; cc += 1024){
			double u = 0.0;
			if(cc != g){
				u = M[(size_t)cc * n + g] * inv;
				M[(size_t)cc * n + g] = u;
			}
			rowg[cc] = u;
		}
		__syncthreads();
		// rank-1 update over all threads: thread (r, c0) with r = tid % n, c0 = tid / n takes the columns
		// c0, c0 + cpp, ... (n <= 1024); for larger n every thread takes rows tid, tid + 1024 of all columns
		{
			const bool small = n <= 1024;
			const unsigned cpp = small ? 1024 / n : 1, rstep = small ? n : 1024;
			const unsigned r0 = small ? tid % n : tid, c0 = small ? tid / n : 0;
			if(c0 < cpp)
				for(unsigned cc = c0; cc < n; cc += cpp){
					double *mc = M + (size_t)cc * n;
					const double u = rowg[cc];
					for(unsigned r = r0; r < n; r += rstep){
						if(r == g) continue;
						if(cc == g) mc[r] = -colg[r] * inv;
						else mc[r] -= colg[r] * u;
					}
				}
		}
		if(tid == 0) M[(size_t)g * n + g] = inv;         // (no thread writes row g in the update)
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

__global__ __launch_bounds__(1024) void gj_fallback_kernel(double *__restrict__ M, unsigned n, double *__restrict__ scal,
		int *__restrict__ piv_out)
{
	gj_fallback_body(M, n, scal, piv_out);
}

// M (column-major, N x N, FP64) <- M^{-1}; log|det| and sign into the device scalars.
int gj_logdet_inverse(ephmc_ctx *c, double *M)
{
	const unsigned n = c->ns;
	if(n > 4096){ set_error("gj: N = %u exceeds 4096", n); return EPHMC_ERR_UNSUPPORTED; }
	const double *cur = M;
	double *buf[2] = {c->d_gj[0], c->d_gj[1]};
	int which = 0;
	// D^{-1} and L of the current and the next diagonal block (double-buffered by step parity)
	double *dinv_g[2] = {c->d_dlu, c->d_dlu + GJ_NB * GJ_NB}, *lm_g[2] = {c->d_dlu + 2 * GJ_NB * GJ_NB, c->d_dlu + 3 * GJ_NB * GJ_NB};
	gj_first_kernel<<<1, 256, 0, c->stream>>>(M, n, std::min<unsigned>(GJ_NB, n), dinv_g[0], lm_g[0], c->d_diag, c->d_piv, c->d_scal);
	EPHMC_LAUNCHED();
	const dim3 sgrid((n + GJ_RB - 1) / GJ_RB + 1, (n + GJ_CB - 1) / GJ_CB);
	int par = 0;
	// from the second step on the input M is no longer read by the steps, so the last step writes
	// the inverse into it and sums the log det itself; a single-step inversion (N <= 32) copies
	const bool direct = n > GJ_NB;
	for(unsigned k0 = 0; k0 < n; k0 += GJ_NB){
		const unsigned nbe = std::min<unsigned>(GJ_NB, n - k0);
		double *next = buf[which];
		double *final_out = (direct && k0 + nbe >= n) ? M : nullptr;
		gj_step_kernel<<<sgrid, 256, 0, c->stream>>>(cur, next, n, k0, nbe, dinv_g[par], lm_g[par], dinv_g[par ^ 1], lm_g[par ^ 1],
			c->d_diag, c->d_piv, c->d_scal, final_out);
		EPHMC_LAUNCHED();
		cur = next;
		which ^= 1;
		par ^= 1;
	}
	if(!direct){
		gj_logdet_kernel<<<1, 1024, 0, c->stream>>>(c->d_diag, n, c->d_scal);
		EPHMC_LAUNCHED();
		gj_copy_kernel<<<1024, 256, 0, c->stream>>>(cur, M, c->mat, c->d_scal);
		EPHMC_LAUNCHED();
	}
	const size_t lds = (size_t)n * (2 * sizeof(double) + sizeof(int));
	(void)hipFuncSetAttribute((const void *)gj_fallback_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
	gj_fallback_kernel<<<1, 1024, lds, c->stream>>>(M, n, c->d_scal, c->d_piv);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// ---- many chains in one launch (batch.hip): the kernels of gj_logdet_inverse with the arguments
// of chain blockIdx.z read from a table; a chain's blocks past its own grid return.  Only the
// multi-step form (N > 32, the inverse written by the last step) is batched.
struct GjFirstArgs {
	const double *A;
	unsigned n, nbe;
	double *dinv_out, *lm_out, *diag;
	int *piv;
	double *scal;
};

__global__ __launch_bounds__(256) void gj_first_multi_kernel(const GjFirstArgs *__restrict__ tab)
{
	const GjFirstArgs &a = tab[blockIdx.z];
	gj_first_body(a.A, a.n, a.nbe, a.dinv_out, a.lm_out, a.diag, a.piv, a.scal);
}

struct GjStepArgs {
	const double *A;
	double *An;
	unsigned n, k0, nbe;
	const double *dinv_cur, *lm_cur;
	double *dinv_next, *lm_next, *diag;
	int *piv;
	double *scal, *out_final;
};

__global__ __launch_bounds__(256) void gj_step_multi_kernel(const GjStepArgs *__restrict__ tab)
{
	const GjStepArgs &a = tab[blockIdx.z];
	const unsigned nrb = (a.n + GJ_RB - 1) / GJ_RB, ncb = (a.n + GJ_CB - 1) / GJ_CB;
	if(blockIdx.x > nrb || blockIdx.y >= ncb) return;
	gj_step_body(a.A, a.An, a.n, a.k0, a.nbe, a.dinv_cur, a.lm_cur, a.dinv_next, a.lm_next, a.diag, a.piv, a.scal, a.out_final,
		blockIdx.x, blockIdx.y, ncb);
}

struct GjFallbackArgs {
	double *M;
	unsigned n;
	double *scal;
	int *piv_out;
};

__global__ __launch_bounds__(1024) void gj_fallback_multi_kernel(const GjFallbackArgs *__restrict__ tab)
{
	const GjFallbackArgs &a = tab[blockIdx.z];
	gj_fallback_body(a.M, a.n, a.scal, a.piv_out);
}

int batch_gj(BatchOp &first, std::vector<BatchOp> &steps, BatchOp &fallback, ephmc_ctx *c, double *M)
{
	const unsigned n = c->ns;
	if(n <= GJ_NB || n > 4096){ set_error("batch gj: N = %u (the batch inverts 33 <= N <= 4096)", n); return EPHMC_ERR_UNSUPPORTED; }
	double *dinv_g[2] = {c->d_dlu, c->d_dlu + GJ_NB * GJ_NB}, *lm_g[2] = {c->d_dlu + 2 * GJ_NB * GJ_NB, c->d_dlu + 3 * GJ_NB * GJ_NB};
	const GjFirstArgs fa{M, n, (unsigned)GJ_NB, dinv_g[0], lm_g[0], c->d_diag, c->d_piv, c->d_scal};
	EPHMC_TRY(first.add((const void *)gj_first_multi_kernel, 256, 0, 1, 1, fa));
	const unsigned nrb = (n + GJ_RB - 1) / GJ_RB, ncb = (n + GJ_CB - 1) / GJ_CB;
	const unsigned nsteps = (n + GJ_NB - 1) / GJ_NB;
	if(steps.size() < nsteps) steps.resize(nsteps);
	const double *cur = M;
	double *buf[2] = {c->d_gj[0], c->d_gj[1]};
	int which = 0, par = 0;
	for(unsigned i = 0; i < nsteps; i++){
		const unsigned k0 = i * GJ_NB, nbe = std::min<unsigned>(GJ_NB, n - k0);
		double *next = buf[which];
		GjStepArgs a{cur, next, n, k0, nbe, dinv_g[par], lm_g[par], dinv_g[par ^ 1], lm_g[par ^ 1], c->d_diag, c->d_piv,
			c->d_scal, k0 + nbe >= n ? M : nullptr};
		EPHMC_TRY(steps[i].add((const void *)gj_step_multi_kernel, 256, 0, nrb + 1, ncb, a));
		cur = next;
		which ^= 1;
		par ^= 1;
	}
	const size_t lds = (size_t)n * (2 * sizeof(double) + sizeof(int));
	(void)hipFuncSetAttribute((const void *)gj_fallback_multi_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
		(int)std::max(lds, fallback.lds));
	const GjFallbackArgs fb{M, n, c->d_scal, c->d_piv};
	return fallback.add((const void *)gj_fallback_multi_kernel, 1024, lds, 1, 1, fb);
}

} // namespace ephmc
