// tl.hip -- transient localisation with static disorder on the GPU (FP64), replacing
// hist_trans_loc x meas (organic_trans_loc.c:25-83, :101-105).
//
// Per disorder sample s (3N normals x_s, drawn by the host in the reference's order):
//   H = -mu + K(x_s)                     (construct_hamilton with -mu, organic_dense_matrix.c:218-224)
//   H = V diag(E) V^T                    (LAPACKE_dsyev, organic_trans_loc.c:40)
//   for all pairs E_j >= E_i:            (organic_trans_loc.c:45-78)
//     proj_a = v_j . I_a v_i,  I_nm = sum_c (sum_a C[c,a] proj_a)^2 = sum_c (v_j . G_c v_i)^2,
//     G_c = sum_a C[c,a] I_a (antisymmetric bond currents, apply_current_to_v :226-249),
//     greens[(unsigned)((E_j - E_i)/width + .5)] += (e^{-E_i} + e^{-E_j}) I_nm / (width N)
//   results[s] = sum_i e^{-E_i} / N
//
// Kernels, a batch of B samples at a time (one workgroup per sample unless stated):
//   tl_build        dense H (column c contiguous; H is symmetric)
//   tl_tridiag      Householder reduction to tridiagonal form (lower) on the full symmetric
//                   matrix, in panels of 8 reflectors (N <= 1024): each step reads the trailing
//                   block once (its matrix-vector product, corrected by the panel's V and W) and
//                   one rank-16 update per panel writes it.  Larger N: one column at a time, the
//                   rank-2 update of step j fused with the matrix-vector product of step j+1
//                   (read and written once per step).  Reflectors v_j overwrite column j below
//                   the diagonal.
//   tl_ql_scalar    implicit QL with Wilkinson shifts on (d, e) (the scalar recurrence of the
//                   EISPACK tql2 method, one wave per sample computing it in every lane),
//                   recording every rotation (c, s) and the position range of every sweep
//   tl_ql_apply     the recorded rotations applied to Z = I, thread = row of Z: G sweeps per
//                   pass with a register window of G+1 columns (a wavefront: sweep g runs one
//                   position behind sweep g-1), so Z is read and written once per G sweeps;
//                   columns are loaded several steps ahead
//   tl_back         V = H_0 H_1 ... H_{N-3} Z in compact-WY blocks of 32 reflectors (N <= 1024):
//                   tl_yb (each block's reflectors as a row-major N x 32 panel) and tl_wy_t (the
//                   blocks' T factors), then
//                   tl_wy_apply, a workgroup per 16 columns of Z held in matrix-core accumulators
//                   for the whole sweep (Z <- Z - Y (T (Y^T Z)) per block, last block first);
//                   larger N: a wave holds CW columns of Z in registers and applies every
//                   reflector (last first)
//   tl_sort         eigenvalue ranks, e^{-E}, the partition sum
//   tl_current      V sorted by eigenvalue and W_c = G_c V (gather form of the bond currents)
//   tl_pairs        64 x 64 tiles of J_c = V^T W_c (FP64 matrix cores, v_mfma_f64_16x16x4)
//                   with the binning in the epilogue; tiles that cannot hold a pair E_j >= E_i are skipped.  Each tile
//                   bins its 4096 pairs in a fixed order (no atomics), so bins are
//                   reproducible run to run
//   tl_reduce       fixed-order sum of the tile histograms into the accumulated bins
// Pairs whose bin index is >= nbins are dropped (the reference writes them past the end of its
// greens array).
#include "internal.h"
#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace ephmc {

constexpr int TL_BT = 1024;     // tridiagonalisation / sort workgroup
constexpr int TL_G = 24;        // sweeps per wavefront pass
constexpr int TL_PF = 8;        // columns loaded ahead by the rotation pass
constexpr int TL_QC = 66;       // rotation steps staged in LDS at a time (a multiple of TL_G + 1 + TL_PF)
constexpr int TL_TILE = 64;
constexpr int TL_KC = 16;       // rows of V / W per LDS stage of the pair tiles
constexpr unsigned TL_HB = 512; // bins a pair tile accumulates in LDS (wider tiles bin by search)
constexpr unsigned TL_NONE = 0xffffffffu;

__device__ inline double wave_sum(double v)
{
#pragma unroll
	for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;          // identical in every lane (a + b == b + a)
}

// deterministic block sum; every thread returns the same value
template <int BT>
__device__ double block_sum(double v, double *red)
{
	v = wave_sum(v);
	if((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
	__syncthreads();
	double t = 0;
#pragma unroll
	for(int q = 0; q < BT / 64; q++) t += red[q];
	__syncthreads();
	return t;
}

// ---- H ----------------------------------------------------------------------------------
__global__ void tl_build_kernel(const double *__restrict__ xs, const int *__restrict__ nb, const unsigned *__restrict__ link,
		double t0, double t1, double t2, double k0, double k1, double k2, double diag, unsigned n, double *__restrict__ Aall)
{
	const unsigned s = blockIdx.y;
	const double *x = xs + (size_t)s * 3 * n;
	double *A = Aall + (size_t)s * n * n;
	const size_t nn = (size_t)n * n;
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		double h = r == c ? diag : 0.0;
#pragma unroll
		for(int d = 0; d < 6; d++){
			if((unsigned)nb[d * n + r] == c){
				const int k = d >> 1;
				const double tk = k == 0 ? t0 : (k == 1 ? t1 : t2), kk = k == 0 ? k0 : (k == 1 ? k1 : k2);
				h += tk - kk * x[link[d * n + r]];
			}
		}
		A[e] = h;
	}
}

// ---- Householder tridiagonalisation --------------------------------------------------------
// x = A[j][j+1 .. n-1] in vb; makes v (v[j+1] = 1) in vb and in column j of A, returns tau.
template <int BT, int RPT>
__device__ double tl_reflector(double *A, unsigned n, unsigned j, double *vb, double *red, double *e, double *tau)
{
	double ss = 0;
#pragma unroll
	for(int q = 0; q < RPT; q++){
		const unsigned r = threadIdx.x + q * BT;
		if(r >= j + 2 && r < n) ss += vb[r] * vb[r];
	}
	ss = block_sum<BT>(ss, red);
	const double alpha = vb[j + 1];
	double tj = 0, beta = alpha, scale = 0;
	if(ss != 0){
		beta = -copysign(hypot(alpha, sqrt(ss)), alpha);
		tj = (beta - alpha) / beta;
		scale = 1.0 / (alpha - beta);
	}
#pragma unroll
	for(int q = 0; q < RPT; q++){
		const unsigned r = threadIdx.x + q * BT;
		if(r >= j + 1 && r < n){
			const double v = r == j + 1 ? 1.0 : vb[r] * scale;
			vb[r] = v;
			A[(size_t)j * n + r] = v;
		}
	}
	if(threadIdx.x == 0){
		e[j] = beta;
		tau[j] = tj;
	}
	__syncthreads();
	return tj;
}

template <int BT, int RPT>
__global__ __launch_bounds__(BT) void tl_tridiag_kernel(double *__restrict__ Aall, unsigned n, double *__restrict__ dall,
		double *__restrict__ eall, double *__restrict__ tall)
{
	extern __shared__ double sh[];
	double *vA = sh, *vB = sh + n, *w = sh + 2 * n, *red = sh + 3 * n;
	const unsigned s = blockIdx.x, tid = threadIdx.x;
	double *A = Aall + (size_t)s * n * n, *d = dall + (size_t)s * n, *e = eall + (size_t)s * n, *tau = tall + (size_t)s * n;
	if(n <= 2){
		if(tid == 0){
			d[0] = A[0];
			e[0] = 0;
			tau[0] = 0;
			if(n == 2){
				d[1] = A[3];
				e[0] = A[1];
				e[1] = 0;
				tau[1] = 0;
			}
		}
		return;
	}
	double p[RPT];
	// step 0: reflector of column 0 and p = tau A22 v
#pragma unroll
	for(int q = 0; q < RPT; q++){
		const unsigned r = tid + q * BT;
		if(r >= 1 && r < n) vA[r] = A[r];
	}
	if(tid == 0) d[0] = A[0];
	double tj = tl_reflector<BT, RPT>(A, n, 0, vA, red, e, tau);
	double kk = 0;
#pragma unroll
	for(int q = 0; q < RPT; q++){
		const unsigned r = tid + q * BT;
		p[q] = 0;
		if(r >= 1 && r < n){
			double acc = 0;
			for(unsigned c = 1; c < n; c++) acc += A[(size_t)c * n + r] * vA[c];
			p[q] = tj * acc;
			kk += p[q] * vA[r];
		}
	}
	kk = 0.5 * tj * block_sum<BT>(kk, red);
#pragma unroll
	for(int q = 0; q < RPT; q++){
		const unsigned r = tid + q * BT;
		if(r >= 1 && r < n) w[r] = p[q] - kk * vA[r];
	}
	__syncthreads();
	double *v = vA, *vn = vB;
	for(unsigned j = 0; j + 3 < n; j++){
		// (v, w) of step j act on rows/columns j+1 .. n-1.  Column j+1 first:
		const double wj1 = w[j + 1], vj1 = v[j + 1];
#pragma unroll
		for(int q = 0; q < RPT; q++){
			const unsigned r = tid + q * BT;
			if(r >= j + 1 && r < n){
				const size_t o = (size_t)(j + 1) * n + r;
				const double a = A[o] - (v[r] * wj1 + w[r] * vj1);
				A[o] = a;
				if(r == j + 1) d[j + 1] = a;
				else vn[r] = a;
			}
		}
		tj = tl_reflector<BT, RPT>(A, n, j + 1, vn, red, e, tau);
		// rank-2 update of step j on rows/columns j+2 .. n-1, fused with p = tau A22 vn
		kk = 0;
#pragma unroll
		for(int q = 0; q < RPT; q++){
			const unsigned r = tid + q * BT;
			p[q] = 0;
			if(r >= j + 2 && r < n){
				const double vr = v[r], wr = w[r];
				double acc = 0;
				unsigned c = j + 2;
				for(; c + 4 <= n; c += 4){
					const size_t o = (size_t)c * n + r;
					const double a0 = A[o] - (vr * w[c] + wr * v[c]);
					const double a1 = A[o + n] - (vr * w[c + 1] + wr * v[c + 1]);
					const double a2 = A[o + 2 * (size_t)n] - (vr * w[c + 2] + wr * v[c + 2]);
					const double a3 = A[o + 3 * (size_t)n] - (vr * w[c + 3] + wr * v[c + 3]);
					A[o] = a0;
					A[o + n] = a1;
					A[o + 2 * (size_t)n] = a2;
					A[o + 3 * (size_t)n] = a3;
					acc += a0 * vn[c];
					acc += a1 * vn[c + 1];
					acc += a2 * vn[c + 2];
					acc += a3 * vn[c + 3];
				}
				for(; c < n; c++){
					const size_t o = (size_t)c * n + r;
					const double a = A[o] - (vr * w[c] + wr * v[c]);
					A[o] = a;
					acc += a * vn[c];
				}
				p[q] = tj * acc;
				kk += p[q] * vn[r];
			}
		}
		kk = 0.5 * tj * block_sum<BT>(kk, red);
#pragma unroll
		for(int q = 0; q < RPT; q++){
			const unsigned r = tid + q * BT;
			if(r >= j + 2 && r < n) w[r] = p[q] - kk * vn[r];
		}
		double *t = v;
		v = vn;
		vn = t;
		__syncthreads();
	}
	// (v, w) of step n-3 on the last 2 x 2 block
	if(tid == 0){
		const unsigned a = n - 2, b = n - 1;
		d[a] = A[(size_t)a * n + a] - 2 * v[a] * w[a];
		e[a] = A[(size_t)a * n + b] - (v[b] * w[a] + w[b] * v[a]);
		d[b] = A[(size_t)b * n + b] - 2 * v[b] * w[b];
		e[b] = 0;
		tau[a] = 0;
		tau[b] = 0;
	}
}

// Panel form of the same reduction (the scheme of LAPACK's dlatrd): for the reflectors
// j = j0 .. j0 + NB - 1 the trailing matrix is left as it was at the start of the panel, and the
// updates of the panel's earlier steps enter through their (v_i, w_i):
//     column j:   a = A[:, j] - sum_i (v_i w_i[j] + w_i v_i[j]),
//     product:    A' v = A v - V (W^T v) - W (V^T v),     w = tau A' v - (tau^2 / 2)(v^T A' v) v,
// so each step reads the trailing block once and writes nothing; one update
// A -= V W^T + W V^T per panel writes it.  Only the lower triangle (row >= column) is read and
// written: wave w owns rows 64w .. 64w + 63 and walks its row block's lower part in tiles of 64 x 8;
// lane r adds A[r][c] v[c] to its row's product and A[r][c] v[r] goes to column c's, whose 8 sums
// over the wave's rows a butterfly reduce-scatter forms (10 lane exchanges per tile) and an LDS
// triangle of per-wave column sums holds; each row then adds the column sums of the waves below it
// in wave order, a fixed order.  Thread r keeps row r of the panel's V and W in registers
// (v_i[r], w_i[r]); row j of them goes through LDS when column j needs it, and 64 rows at a time
// when the panel's update needs them.  Reflectors, d, e and tau come out as from tl_tridiag_kernel.
constexpr int TL_PNB = 8;        // reflectors per panel

// LDS doubles of the panel kernel for n rows on BT threads: vb, block sums, row j of V and W, wave
// dot partials, their sums, the update's rows of V and W, and the triangle of per-wave column sums
template <int BT, int NB>
constexpr size_t tl_panel_lds_doubles(unsigned n)
{
	constexpr size_t NW = BT / 64;
	return (size_t)n + NW + 2 * NB + NW * 2 * NB + 2 * NB + 64 * 2 * NB + 64 * NW * (NW + 1) / 2 + 64 * NW;
}

// Sums over the 64 lanes of a wave of x[0..7] (one value per lane each): returns, in lane l, the
// sum of x[l >> 3] (the same in the eight lanes that differ from l in bits 0..2).  Butterfly
// reduce-scatter: at each exchange a lane keeps half of its columns and adds its partner's half,
// in a fixed order.
__device__ __forceinline__ double tl_reduce8(const double (&x)[8])
{
	const unsigned lane = threadIdx.x & 63;
	double a[4], b[2], d;
	const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
#pragma unroll
	for(int i = 0; i < 4; i++){
		const double send = h5 ? x[i] : x[4 + i];
		const double keep = h5 ? x[4 + i] : x[i];
		a[i] = keep + __shfl_xor(send, 32, 64);
	}
#pragma unroll
	for(int i = 0; i < 2; i++){
		const double send = h4 ? a[i] : a[2 + i];
		const double keep = h4 ? a[2 + i] : a[i];
		b[i] = keep + __shfl_xor(send, 16, 64);
	}
	{
		const double send = h3 ? b[0] : b[1];
		const double keep = h3 ? b[1] : b[0];
		d = keep + __shfl_xor(send, 8, 64);
	}
	d += __shfl_xor(d, 4, 64);
	d += __shfl_xor(d, 2, 64);
	d += __shfl_xor(d, 1, 64);
	return d;
}

template <int BT, int RPT, int NB>
__global__ __launch_bounds__(BT) void tl_tridiag_panel_kernel(double *__restrict__ Aall, unsigned n, double *__restrict__ dall,
		double *__restrict__ eall, double *__restrict__ tall)
{
	extern __shared__ double sh[];
	constexpr int NW = BT / 64;
	double *vb = sh;                         // n: column j, then v_j
	double *red = vb + n;                    // NW
	double *rowvw = red + NW;                // 2 NB: v_i[j], w_i[j]
	double *dots = rowvw + 2 * NB;           // NW x 2 NB wave partials
	double *dsum = dots + NW * 2 * NB;       // 2 NB: W^T v, V^T v
	double *chunk = dsum + 2 * NB;           // 64 x 2 NB: rows of V and W for the panel update
	double *zsum = chunk + 64 * 2 * NB;      // row block I's column sums at 64 I (I + 1) / 2 + column
	double *rowpart = zsum + 64 * NW * (NW + 1) / 2;   // row sums of block I's tiles done by its partner wave
	static_assert(RPT == 1, "the lower-triangle product holds one row per thread");
	const unsigned s = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	double *A = Aall + (size_t)s * n * n, *d = dall + (size_t)s * n, *e = eall + (size_t)s * n, *tau = tall + (size_t)s * n;
	if(n <= 2){
		if(tid == 0){
			d[0] = A[0];
			e[0] = 0;
			tau[0] = 0;
			if(n == 2){
				d[1] = A[3];
				e[0] = A[1];
				e[1] = 0;
				tau[1] = 0;
			}
		}
		return;
	}
	double vrow[RPT][NB], wrow[RPT][NB];
	for(unsigned j0 = 0; j0 + 2 < n; j0 += NB){
		const unsigned kmax = min((unsigned)NB, n - 2 - j0), jn = j0 + kmax;
#pragma unroll
		for(int q = 0; q < RPT; q++)
#pragma unroll
			for(int i = 0; i < NB; i++){ vrow[q][i] = 0; wrow[q][i] = 0; }
		for(unsigned k = 0; k < kmax; k++){
			const unsigned j = j0 + k;
			// row j of V and W
#pragma unroll
			for(int q = 0; q < RPT; q++)
				if(tid + q * BT == j){
#pragma unroll
					for(int i = 0; i < NB; i++){ rowvw[i] = vrow[q][i]; rowvw[NB + i] = wrow[q][i]; }
				}
			__syncthreads();
			// column j of the updated matrix, rows j .. n-1
#pragma unroll
			for(int q = 0; q < RPT; q++){
				const unsigned r = tid + q * BT;
				if(r >= j && r < n){
					double a = A[(size_t)j * n + r];
#pragma unroll
					for(int i = 0; i < NB; i++) a -= vrow[q][i] * rowvw[NB + i] + wrow[q][i] * rowvw[i];
					if(r == j) d[j] = a;
					else vb[r] = a;
				}
			}
			__syncthreads();
			const double tj = tl_reflector<BT, RPT>(A, n, j, vb, red, e, tau);
			// A v over the panel-start trailing block from its lower triangle, and the panel's
			// W^T v, V^T v.  Row block I (64 rows) has one tile of 64 x 8 per 8 columns from the
			// chunk of column j + 1 to its diagonal, so blocks far down hold more tiles; waves are
			// paired (w, NW-1-w) and the pair's tiles split evenly: wave w < NW/2 takes its own
			// block's tiles and the partner block's first ones, whose row sums it leaves in LDS
			// (rowpart) for the partner to add after its own.  Column sums go to the triangle
			// zsum by row block.  Lanes past n or above j + 1 load nothing and add zeros.
			double acc[RPT];
			{
				const unsigned cstart = (j + 1) & ~7u;
				auto chunks_of = [&](unsigned I) -> unsigned {
					const unsigned cend = 64 * I + 64;
					return (64 * I < n && cend > cstart) ? (cend - cstart) / 8 : 0u;
				};
				// (the wave index as a wave-uniform value, so that the tile bookkeeping stays scalar)
				const unsigned wv = (unsigned)__builtin_amdgcn_readfirstlane((int)wave);
				const unsigned partner = NW - 1 - wv;
				// the lower block of the pair, its chunks, and the upper block's first chunks that the
				// lower block's wave takes
				const unsigned lo_blk = wv < NW / 2 ? wv : partner, hi_blk = NW - 1 - lo_blk;
				const unsigned c_lo = chunks_of(lo_blk), c_hi = chunks_of(hi_blk);
				const unsigned t_lo = (c_lo + c_hi) / 2;
				const unsigned lent = t_lo > c_lo ? t_lo - c_lo : 0u;
				// a run of `cnt` tiles of row block I from chunk `first`: row sums into s0, s1
				auto run = [&](unsigned I, unsigned first, unsigned cnt, double &s0, double &s1){
					const unsigned r = 64 * I + lane;
					const bool live = r >= j + 1 && r < n;
					const double vr = live ? vb[r] : 0.0;
					double *zw = zsum + 32 * (size_t)I * (I + 1);
					const unsigned cbeg = cstart + 8 * first, cend = cbeg + 8 * cnt;
					for(unsigned c0 = cbeg; c0 < cend; c0 += 8){
						double x[8];
#pragma unroll
						for(int k = 0; k < 8; k++){
							const unsigned c = c0 + k;
							x[k] = (live && c >= j + 1 && c <= r) ? A[(size_t)c * n + r] : 0.0;
						}
						// (v outside [j + 1, n) is not read: those entries of vb are stale or past its end)
#pragma unroll
						for(int k = 0; k < 8; k += 2){
							const unsigned c = c0 + k;
							s0 = fma(x[k], (c >= j + 1 && c < n) ? vb[c] : 0.0, s0);
							s1 = fma(x[k + 1], (c + 1 >= j + 1 && c + 1 < n) ? vb[c + 1] : 0.0, s1);
						}
						// then the column part, in place
#pragma unroll
						for(int k = 0; k < 8; k++) x[k] = c0 + k < r ? x[k] * vr : 0.0;
						const double zc = tl_reduce8(x);
						if((lane & 7) == 0) zw[c0 + (lane >> 3)] = zc;
					}
				};
				// one or two runs through one copy of the tile loop: the wave's own block, then (lower
				// waves) the partner block's first `lent` tiles
				double own = 0;
#pragma nounroll
				for(int seg = 0; seg < (wv < NW / 2 ? 2 : 1); seg++){
					const unsigned I = seg ? partner : wv;
					const unsigned first = wv < NW / 2 ? 0u : lent;
					const unsigned cnt = seg ? lent : (wv < NW / 2 ? c_lo : c_hi - lent);
					double s0 = 0, s1 = 0;
					run(I, first, cnt, s0, s1);
					if(seg) rowpart[partner * 64 + lane] = s0 + s1;
					else own = s0 + s1;
				}
				acc[0] = own;
				__syncthreads();
				const unsigned r = tid;
				const bool live = r >= j + 1 && r < n;
				if(wv >= NW / 2 && lent) acc[0] += rowpart[wv * 64 + lane];
				// the column sums of the rows below r: row blocks r / 64 .. the last with rows < n
				if(live)
					for(unsigned w2 = r >> 6; 64 * w2 < n; w2++) acc[0] += zsum[32 * (size_t)w2 * (w2 + 1) + r];
			}
			double pw[NB], pv[NB];
#pragma unroll
			for(int i = 0; i < NB; i++){ pw[i] = 0; pv[i] = 0; }
			if(tid >= j + 1 && tid < n){
				const double vr = vb[tid];
#pragma unroll
				for(int i = 0; i < NB; i++){
					pw[i] = fma(wrow[0][i], vr, pw[i]);
					pv[i] = fma(vrow[0][i], vr, pv[i]);
				}
			}
			if(k > 0){
				// block sums of the 2k dot products (fixed order: waves, then their partials in order)
#pragma unroll
				for(int i = 0; i < NB; i++)
					if(i < (int)k){
						pw[i] = wave_sum(pw[i]);
						pv[i] = wave_sum(pv[i]);
					}
				if(lane == 0)
#pragma unroll
					for(int i = 0; i < NB; i++)
						if(i < (int)k){
							dots[wave * 2 * NB + i] = pw[i];
							dots[wave * 2 * NB + NB + i] = pv[i];
						}
				__syncthreads();
				if(tid < 2 * NB){
					double t = 0;
					for(int w = 0; w < NW; w++) t += dots[w * 2 * NB + tid];
					dsum[tid] = t;
				}
				__syncthreads();
			}
			// p = tau (A v - V (W^T v) - W (V^T v)), kk = tau / 2 p.v, w = p - kk v
			double p[RPT], kk = 0;
#pragma unroll
			for(int q = 0; q < RPT; q++){
				const unsigned r = tid + q * BT;
				p[q] = 0;
				if(r >= j + 1 && r < n){
					double x = acc[q];
#pragma unroll
					for(int i = 0; i < NB; i++)
						if(i < (int)k) x -= vrow[q][i] * dsum[i] + wrow[q][i] * dsum[NB + i];
					p[q] = tj * x;
					kk += p[q] * vb[r];
				}
			}
			kk = 0.5 * tj * block_sum<BT>(kk, red);
#pragma unroll
			for(int q = 0; q < RPT; q++){
				const unsigned r = tid + q * BT;
				const bool in = r >= j + 1 && r < n;
				const double vr = in ? vb[r] : 0.0, wr = in ? p[q] - kk * vr : 0.0;
#pragma unroll
				for(int i = 0; i < NB; i++)
					if(i == (int)k){ vrow[q][i] = vr; wrow[q][i] = wr; }
			}
			__syncthreads();          // vb is rewritten by the next column
		}
		// the panel's update of the trailing block, rows and columns jn .. n-1
		for(unsigned c0 = jn; c0 < n; c0 += 64){
#pragma unroll
			for(int q = 0; q < RPT; q++){
				const unsigned r = tid + q * BT;
				if(r >= c0 && r < c0 + 64 && r < n){
#pragma unroll
					for(int i = 0; i < NB; i++){
						chunk[(r - c0) * 2 * NB + i] = vrow[q][i];
						chunk[(r - c0) * 2 * NB + NB + i] = wrow[q][i];
					}
				}
			}
			__syncthreads();
			const unsigned cn = min(64u, n - c0);
#pragma unroll
			for(int q = 0; q < RPT; q++){
				const unsigned r = tid + q * BT;
				if(r >= jn && r < n){
					// the lower triangle only (column <= row)
					const unsigned cend = r + 1 - c0 < cn ? r + 1 - c0 : cn;
					for(unsigned cc = 0; cc < cend && r >= c0; cc++){
						const double *cv = chunk + cc * 2 * NB;
						double x = 0;
#pragma unroll
						for(int i = 0; i < NB; i++) x += vrow[q][i] * cv[NB + i] + wrow[q][i] * cv[i];
						A[(size_t)(c0 + cc) * n + r] -= x;
					}
				}
			}
			__syncthreads();
		}
	}
	// the last 2 x 2 block, already updated
	if(tid == 0){
		const unsigned a = n - 2, b = n - 1;
		d[a] = A[(size_t)a * n + a];
		e[a] = A[(size_t)a * n + b];
		d[b] = A[(size_t)b * n + b];
		e[b] = 0;
		tau[a] = 0;
		tau[b] = 0;
	}
}

// ---- implicit QL: scalar recurrence ----------------------------------------------------------
// e[i] couples i and i+1.  Rotation (c, s) at position i acts on columns (i, i+1) of Z:
// Z(:,i+1) <- s Z(:,i) + c Z(:,i+1),  Z(:,i) <- c Z(:,i) - s Z(:,i+1).  Sweep records: (first
// rotation index, lowest position, highest position); positions run from high to low.
// One wave per sample runs the recurrence with every lane computing the same values (d, e in
// LDS, written by all lanes alike); the deflation test scans 64 positions per step.

// 1/x: hardware reciprocal refined by two Newton steps
__device__ inline double rcp_nr(double x)
{
	double r = __builtin_amdgcn_rcp(x);
	double t = fma(-x, r, 1.0);
	r = fma(r, t, r);
	t = fma(-x, r, 1.0);
	return fma(r, t, r);
}

// sqrt(f^2 + g^2): the plain form while the square sum is a normal number, hypot otherwise
__device__ inline double norm2(double f, double g)
{
	const double h2 = fma(f, f, g * g);
	if(h2 > 0x1p-1000 && h2 < 0x1p1000) return sqrt(h2);
	return hypot(f, g);
}

__global__ __launch_bounds__(64) void tl_ql_scalar_kernel(const double *__restrict__ dall, const double *__restrict__ eall,
		unsigned n, double *__restrict__ evall, double2 *__restrict__ rotall, size_t rot_cap, int4 *__restrict__ swall,
		unsigned sw_cap, unsigned *__restrict__ nsw_all, int *__restrict__ status)
{
	extern __shared__ double sh[];
	double *d = sh, *e = sh + n;
	const unsigned s = blockIdx.x, lane = threadIdx.x;
	for(unsigned i = lane; i < n; i += 64){
		d[i] = dall[(size_t)s * n + i];
		e[i] = i + 1 < n ? eall[(size_t)s * n + i] : 0.0;
	}
	__syncthreads();
	double2 *rot = rotall + (size_t)s * rot_cap;
	int4 *sw = swall + (size_t)s * sw_cap;
	const double eps = 2.220446049250313e-16;
	size_t nrot = 0;
	unsigned nsw = 0;
	int st = 0;
	for(unsigned l = 0; l < n && !st; l++){
		int it = 0;
		for(;;){
			// first m >= l with |e[m]| <= eps (|d[m]| + |d[m+1]|), else n - 1
			unsigned m = n - 1;
			for(unsigned m0 = l; m0 + 1 < n; m0 += 64){
				const unsigned mm = m0 + lane;
				const bool hit = mm + 1 < n && fabs(e[mm]) <= eps * (fabs(d[mm]) + fabs(d[mm + 1]));
				const unsigned long long bal = __ballot(hit);
				if(bal){
					m = m0 + (unsigned)__ffsll(bal) - 1;
					break;
				}
			}
			if(m == l) break;
			if(++it > 60){ st = 1; break; }
			if(nrot + (m - l) > rot_cap || nsw >= sw_cap){ st = 2; break; }
			double g = (d[l + 1] - d[l]) / (2 * e[l]);
			double r = hypot(g, 1.0);
			g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
			double sn = 1, cs = 1, p = 0;
			const size_t base = nrot;
			int i = (int)m - 1;
			bool broke = false;
			// d[i+1], d[i], e[i] are not yet modified by this sweep when position i is reached;
			// position i-1's values are loaded one step ahead
			double di1 = d[m], di = d[i], ei = e[i];
			for(; i >= (int)l; i--){
				const int ip = i > 0 ? i - 1 : 0;
				const double dn = d[ip], en = e[ip];
				const double f = sn * ei, b = cs * ei;
				r = norm2(f, g);
				e[i + 1] = r;
				if(r == 0){
					d[i + 1] = di1 - p;
					e[m] = 0;
					broke = true;
					break;
				}
				const double ir = rcp_nr(r);
				sn = f * ir;
				cs = g * ir;
				g = di1 - p;
				r = (di - g) * sn + 2 * cs * b;
				p = sn * r;
				d[i + 1] = g + p;
				g = cs * r - b;
				if(lane == 0) rot[nrot] = make_double2(cs, sn);
				nrot++;
				di1 = di;
				di = dn;
				ei = en;
			}
			const int hi = (int)m - 1, lo = broke ? i + 1 : (int)l;
			if(lo <= hi){
				if(lane == 0) sw[nsw] = make_int4((int)base, lo, hi, 0);
				nsw++;
			}
			if(broke) continue;
			d[l] -= p;
			e[l] = g;
			e[m] = 0;
		}
	}
	for(unsigned i = lane; i < n; i += 64) evall[(size_t)s * n + i] = d[i];
	if(lane == 0){
		nsw_all[s] = nsw;
		status[s] = st;
	}
}

// ---- rotations applied to Z = I (thread = row, column-major Z) ------------------------------
// G sweeps per pass: sweep k acts at position q + k while the window steps q downwards, so the
// pass holds G+1 columns in registers, and PF more columns are loaded ahead.  The G+1+PF columns
// live in a ring of registers whose slot for logical column k at step t is (k - t) mod (G+1+PF):
// the step loop is unrolled by the ring length, so every slot index is a constant and the window
// moves without register copies (the column leaving the window is stored and its slot takes the
// next column loaded ahead).  The rotations of the next QC steps (a multiple of the ring length)
// of all G sweeps are staged in LDS by the whole workgroup (every row of the sample applies the
// same rotations); a sweep outside its range at a step gets the identity rotation (c, s) = (1, 0),
// which leaves both columns unchanged.
// One step at ring phase PH: the G rotations of the step on the window, then either the window
// written out (the last step, q = qend: returns true) or column q + G stored and its slot loaded
// with column q - 1 - PF.
template <int G, int PF, int PH>
__device__ __forceinline__ bool tl_ql_step(double (&R)[G + 1 + PF], const double2 *rst, int q, int qend, double *__restrict__ Z,
		unsigned n, unsigned r, bool act)
{
	constexpr int NR = G + 1 + PF;
#pragma unroll
	for(int k = 0; k < G; k++){
		const double2 cs = rst[k];
		double &a = R[(k - PH + 2 * NR) % NR], &f = R[(k + 1 - PH + 2 * NR) % NR];
		const double av = a, fv = f;
		f = cs.y * av + cs.x * fv;
		a = cs.x * av - cs.y * fv;
	}
	if(q <= qend){
#pragma unroll
		for(int k = 0; k <= G; k++){
			const int col = qend + k;
			if(act && col >= 0 && col < (int)n) Z[(size_t)col * n + r] = R[(k - PH + 2 * NR) % NR];
		}
		return true;
	}
	constexpr int out = (G - PH + 2 * NR) % NR;
	const int col = q + G;
	if(act && col < (int)n) Z[(size_t)col * n + r] = R[out];
	const int pcol = q - 1 - PF;
	R[out] = (act && pcol >= 0) ? Z[(size_t)pcol * n + r] : 0.0;
	return false;
}

// Steps t0 + PH .. t0 + NR - 1 of a stage (one turn of the ring from phase PH); true when the
// pass ended in it.
template <int G, int PF, int QC, int PH>
__device__ __forceinline__ bool tl_ql_turn(double (&R)[G + 1 + PF], const double2 (*rs)[G], int t0, int qc, int qend,
		double *__restrict__ Z, unsigned n, unsigned r, bool act)
{
	if constexpr(PH < G + 1 + PF){
		if(tl_ql_step<G, PF, PH>(R, rs[t0 + PH], qc - t0 - PH, qend, Z, n, r, act)) return true;
		return tl_ql_turn<G, PF, QC, PH + 1>(R, rs, t0, qc, qend, Z, n, r, act);
	}else{
		return false;
	}
}

template <int G, int PF, int QC>
__global__ __launch_bounds__(256) void tl_ql_apply_kernel(double *__restrict__ Zall, unsigned n, const double2 *__restrict__ rotall,
		size_t rot_cap, const int4 *__restrict__ swall, unsigned sw_cap, const unsigned *__restrict__ nsw_all)
{
	constexpr int NR = G + 1 + PF;
	static_assert(QC % NR == 0, "a stage is whole turns of the ring");
	__shared__ double2 rs[QC][G];
	const unsigned s = blockIdx.y;
	const unsigned r = blockIdx.x * 256 + threadIdx.x;
	const bool act = r < n;
	double *Z = Zall + (size_t)s * n * n;
	const double2 *rot = rotall + (size_t)s * rot_cap;
	const int4 *sw = swall + (size_t)s * sw_cap;
	if(act)
		for(unsigned c = 0; c < n; c++) Z[(size_t)c * n + r] = c == r ? 1.0 : 0.0;
	const unsigned nsw = nsw_all[s];
	for(unsigned g0 = 0; g0 < nsw; g0 += G){
		int lo[G], hi[G], base[G];
		int P = -0x7fffffff, qend = 0x7fffffff;
#pragma unroll
		for(int k = 0; k < G; k++){
			if(g0 + k < nsw){
				const int4 q = sw[g0 + k];
				base[k] = q.x;
				lo[k] = q.y;
				hi[k] = q.z;
				P = max(P, hi[k] - k);
				qend = min(qend, lo[k] - k);
			}else{
				base[k] = 0;
				lo[k] = 1;
				hi[k] = 0;
			}
		}
		// step 0: window columns P .. P+G in slots 0 .. G, columns P-1-j ahead in slot NR-1-j
		double R[NR];
#pragma unroll
		for(int k = 0; k <= G; k++){
			const int col = P + k;
			R[k] = (act && col < (int)n) ? Z[(size_t)col * n + r] : 0.0;
		}
#pragma unroll
		for(int jj = 0; jj < PF; jj++){
			const int col = P - 1 - jj;
			R[NR - 1 - jj] = (act && col >= 0) ? Z[(size_t)col * n + r] : 0.0;
		}
		// steps q = P, P-1, ..., qend (P >= qend) in stages of QC
		bool done = false;
		for(int qc = P; !done; qc -= QC){
			__syncthreads();                 // the previous stage's rotations are consumed
			for(int e = threadIdx.x; e < QC * G; e += 256){
				const int t = e / G, k = e % G, pos = qc - t + k;
				const bool in = qc - t >= qend && pos >= lo[k] && pos <= hi[k];
				rs[t][k] = in ? rot[base[k] + (hi[k] - pos)] : make_double2(1.0, 0.0);
			}
			__syncthreads();
			for(int u = 0; u < QC / NR && !done; u++)
				done = tl_ql_turn<G, PF, QC, 0>(R, rs, u * NR, qc, qend, Z, n, r, act);
		}
	}
}

// ---- back-transformation V = H_0 ... H_{n-3} Z --------------------------------------------
// 256 threads = 4 waves; wave w owns columns (blockIdx.x * 4 + w) * CW + c, c < CW; lane l
// holds rows l + 64 q, q < RPL.
template <int RPL, int CW>
__global__ __launch_bounds__(256) void tl_back_kernel(const double *__restrict__ Aall, const double *__restrict__ tall,
		double *__restrict__ Zall, unsigned n)
{
	const unsigned s = blockIdx.y, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const unsigned col0 = (blockIdx.x * 4 + wave) * CW;
	const double *A = Aall + (size_t)s * n * n, *tau = tall + (size_t)s * n;
	double *Z = Zall + (size_t)s * n * n;
	double z[CW][RPL];
#pragma unroll
	for(int c = 0; c < CW; c++)
#pragma unroll
		for(int q = 0; q < RPL; q++){
			const unsigned r = lane + 64 * q, col = col0 + c;
			z[c][q] = (col < n && r < n) ? Z[(size_t)col * n + r] : 0.0;
		}
	for(int j = (int)n - 3; j >= 0; j--){
		const double tj = tau[j];
		if(tj == 0) continue;
		const double *vj = A + (size_t)j * n;
		const int qlo = (j + 1) >> 6;
		double v[RPL], part[CW];
#pragma unroll
		for(int c = 0; c < CW; c++) part[c] = 0;
#pragma unroll
		for(int q = 0; q < RPL; q++){
			const unsigned r = lane + 64 * q;
			v[q] = 0;
			if(q >= qlo){
				v[q] = (r >= (unsigned)j + 1 && r < n) ? vj[r] : 0.0;
#pragma unroll
				for(int c = 0; c < CW; c++) part[c] += v[q] * z[c][q];
			}
		}
#pragma unroll
		for(int c = 0; c < CW; c++){
			const double sc = tj * wave_sum(part[c]);
#pragma unroll
			for(int q = 0; q < RPL; q++)
				if(q >= qlo) z[c][q] -= sc * v[q];
		}
	}
#pragma unroll
	for(int c = 0; c < CW; c++)
#pragma unroll
		for(int q = 0; q < RPL; q++){
			const unsigned r = lane + 64 * q, col = col0 + c;
			if(col < n && r < n) Z[(size_t)col * n + r] = z[c][q];
		}
}

// ---- back-transformation in compact-WY blocks on the FP64 matrix cores ---------------------
// The reflectors j = 32 b .. 32 b + 31 (of the N - 2) form H_{32b} ... H_{32b+31} = 1 - Y T Y^T,
// Y = [v_j] (N x 32, v_j zero above row j + 1), T upper triangular (T_ii = tau_i,
// T[0:i, i] = -tau_i T[0:i, 0:i] Y[:, 0:i]^T v_i).  V = Q Z is then Z <- Z - Y (T (Y^T Z)) for
// b = last .. 0.

constexpr int TL_WB = 32;        // reflectors per block
constexpr int TL_WC = 16;        // columns of Z per workgroup
constexpr int TL_WW = 8;         // waves per workgroup
constexpr int TL_WRB = 8;        // 16-row blocks per wave (8 waves: N <= 1024)

// Yb[b][r][k] = v_{32 b + k}[r]: the reflectors of each block as a row-major panel of n16 =
// 16 ceil(N / 16) rows x 32 (zero outside their support and in the padding rows), so that a row's
// 32 entries are contiguous; grid (row tiles of 32, blocks, samples).
__global__ __launch_bounds__(256) void tl_yb_kernel(const double *__restrict__ Aall, unsigned n, unsigned nblk,
		double *__restrict__ Yball)
{
	__shared__ double tile[32][33];
	const unsigned s = blockIdx.z, r0 = blockIdx.x * 32, b = blockIdx.y, j0 = b * TL_WB;
	const double *A = Aall + (size_t)s * n * n;
	const unsigned n16 = (n + 15) & ~15u;
	double *Yb = Yball + ((size_t)s * nblk + b) * n16 * TL_WB;
	const unsigned tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for(unsigned k = ty; k < 32; k += 8){
		const unsigned j = j0 + k, r = r0 + tx;          // column j of A, row r
		tile[k][tx] = (j + 3 <= n && r >= j + 1 && r < n) ? A[(size_t)j * n + r] : 0.0;
	}
	__syncthreads();
	for(unsigned k = ty; k < 32; k += 8){
		const unsigned r = r0 + k;
		if(r < n16) Yb[(size_t)r * TL_WB + tx] = tile[tx][k];
	}
}

// T of every block of every sample: grid (blocks, samples).  The Gram entries v_k^T v_i (k < i)
// are summed over 64-row chunks staged in LDS, then the triangular recurrence runs column by column.
__global__ __launch_bounds__(256) void tl_wy_t_kernel(const double *__restrict__ Aall, const double *__restrict__ tall,
		unsigned n, double *__restrict__ Tall)
{
	__shared__ double ych[64][TL_WB + 1];
	__shared__ double G[TL_WB][TL_WB + 1];
	__shared__ double T[TL_WB][TL_WB + 1];
	const unsigned s = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
	const unsigned jb = b * TL_WB, nref = n - 2;
	const double *A = Aall + (size_t)s * n * n, *tau = tall + (size_t)s * n;
	// pairs (k, i), k < i: p = tid and tid + 256 of the 496
	int pk[2], pi[2];
	for(int h = 0; h < 2; h++){
		int p = (int)tid + 256 * h, i = 1;
		while(p >= i && i < TL_WB){ p -= i; i++; }
		pk[h] = i < TL_WB ? p : -1;
		pi[h] = i;
	}
	double g[2] = {0, 0};
	for(unsigned r0 = jb + 1; r0 < n; r0 += 64){
		for(unsigned e = tid; e < 64 * TL_WB; e += 256){
			const unsigned k = e / 64, rr = e % 64, j = jb + k, r = r0 + rr;
			ych[rr][k] = (j < nref && r >= j + 1 && r < n) ? A[(size_t)j * n + r] : 0.0;
		}
		__syncthreads();
		for(int h = 0; h < 2; h++)
			if(pk[h] >= 0)
				for(int rr = 0; rr < 64; rr++) g[h] = fma(ych[rr][pk[h]], ych[rr][pi[h]], g[h]);
		__syncthreads();
	}
	for(int h = 0; h < 2; h++) if(pk[h] >= 0) G[pk[h]][pi[h]] = g[h];
	for(unsigned e = tid; e < TL_WB * TL_WB; e += 256) T[e / TL_WB][e % TL_WB] = 0.0;
	__syncthreads();
	if(tid == 0) T[0][0] = jb < nref ? tau[jb] : 0.0;
	__syncthreads();
	for(int i = 1; i < TL_WB; i++){
		const unsigned j = jb + i;
		const double ti = j < nref ? tau[j] : 0.0;
		double x = 0;
		if((int)tid < i){
			for(int m = (int)tid; m < i; m++) x = fma(T[tid][m], G[m][i], x);
		}
		__syncthreads();
		if((int)tid < i) T[tid][i] = -ti * x;
		if(tid == 0) T[i][i] = ti;
		__syncthreads();
	}
	double *To = Tall + ((size_t)s * gridDim.x + b) * TL_WB * TL_WB;
	for(unsigned e = tid; e < TL_WB * TL_WB; e += 256) To[e] = T[e / TL_WB][e % TL_WB];
}

// Z <- Q Z for 16 columns per workgroup (grid (N / 16, samples)), N <= 1024.  Wave w of the 8 holds
// rows 128 w .. 128 w + 127 of its columns in the matrix-core accumulator layout (16 x 16 blocks: lane
// l, element e = row 4 e + l / 16, column l % 16) for the whole sweep over the blocks, so Z is read
// and written once.  Per block: W^T = Z^T Y (each wave over its rows; the accumulator elements
// serve directly as the A operand, rows 16 q + 4 e + l / 16 of step (q, e)), the eight waves'
// parts summed in LDS, W2^T = W^T T^T, then Z -= Y W2 (v_mfma_f64_16x16x4 both ways).
__global__ __launch_bounds__(64 * TL_WW) void tl_wy_apply_kernel(const double *__restrict__ Yball, const double *__restrict__ Tall,
		unsigned n, unsigned nblk, double *__restrict__ Zall)
{
	typedef double v4d __attribute__((ext_vector_type(4)));
	__shared__ double Ts[TL_WB][TL_WB + 1];
	__shared__ double wpart[TL_WW][2][64][4];
	__shared__ double Wt[TL_WC][TL_WB + 1];
	__shared__ double W2t[TL_WC][TL_WB + 1];
	const unsigned s = blockIdx.y, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const unsigned g = lane >> 4, m = lane & 15, c0 = blockIdx.x * TL_WC, n16 = (n + 15) & ~15u;
	const double *Tb = Tall + (size_t)s * nblk * TL_WB * TL_WB;
	double *Z = Zall + (size_t)s * n * n;
	const unsigned rw0 = wave * 16 * TL_WRB;                      // first row of this wave
	const unsigned col = c0 + m;
	v4d acc[TL_WRB];
#pragma unroll
	for(int q = 0; q < TL_WRB; q++)
#pragma unroll
		for(int e = 0; e < 4; e++){
			const unsigned r = rw0 + 16 * q + 4 * e + g;
			acc[q][e] = (r < n && col < n) ? Z[(size_t)col * n + r] : 0.0;
		}
	for(int b = (int)nblk - 1; b >= 0; b--){
		const unsigned jb = (unsigned)b * TL_WB;
		for(unsigned e = tid; e < TL_WB * TL_WB; e += 64 * TL_WW) Ts[e / TL_WB][e % TL_WB] = Tb[(size_t)b * TL_WB * TL_WB + e];
		// this block's reflector panel: row r at Yb + r * 32 (rows of one lane at fixed offsets)
		const double *Yb = Yball + ((size_t)s * nblk + b) * n16 * TL_WB;
		const double *y1 = Yb + (size_t)(rw0 + g) * TL_WB + m;          // phase 1: row rw0 + 16 q + 4 e + g, column 16 h + m
		const double *y3 = Yb + (size_t)(rw0 + m) * TL_WB + g;          // phase 3: row rw0 + 16 q + m, column 4 t + g
		// W^T partial over this wave's rows r >= jb + 1 (Y is zero above)
		v4d wt[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
		for(int q = 0; q < TL_WRB; q++){
			const unsigned rb = rw0 + 16 * q;
			if(rb + 16 <= jb + 1 || rb >= n) continue;           // uniform per wave
#pragma unroll
			for(int e = 0; e < 4; e++){
				const double a = acc[q][e];
#pragma unroll
				for(int h = 0; h < 2; h++){
					const double y = y1[(16 * q + 4 * e) * TL_WB + 16 * h]; // zero past row n
					wt[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, y, wt[h], 0, 0, 0);
				}
			}
		}
#pragma unroll
		for(int h = 0; h < 2; h++)
#pragma unroll
			for(int e = 0; e < 4; e++) wpart[wave][h][lane][e] = wt[h][e];
		__syncthreads();
		// W^T[c][k]: output element e of lane l of block h is row (column of Z) 4 e + l / 16, column
		// (reflector) 16 h + l % 16
		for(unsigned e2 = tid; e2 < TL_WC * TL_WB; e2 += 64 * TL_WW){
			const unsigned c = e2 / TL_WB, k = e2 % TL_WB, h = k >> 4, l = ((c & 3) << 4) | (k & 15), e = c >> 2;
			double x = wpart[0][h][l][e];
#pragma unroll
			for(int ww = 1; ww < TL_WW; ww++) x += wpart[ww][h][l][e];
			Wt[c][k] = x;
		}
		__syncthreads();
		// W2^T[c][k] = sum_{mm >= k} W^T[c][mm] T[k][mm]; stored negated for the update
		for(unsigned e2 = tid; e2 < TL_WC * TL_WB; e2 += 64 * TL_WW){
			const unsigned c = e2 / TL_WB, k = e2 % TL_WB;
			double x = 0;
			for(unsigned mm = k; mm < TL_WB; mm++) x = fma(Wt[c][mm], Ts[k][mm], x);
			W2t[c][k] = -x;
		}
		__syncthreads();
		// Z[rows][cols] -= Y[rows][k] W2[k][cols]: A = Y (row 16 q + l % 16, k 4 t + l / 16), B = -W2
		// (k 4 t + l / 16, column l % 16)
#pragma unroll
		for(int q = 0; q < TL_WRB; q++){
			const unsigned rb = rw0 + 16 * q;
			if(rb + 16 <= jb + 1 || rb >= n) continue;
#pragma unroll
			for(int t = 0; t < TL_WB / 4; t++){
				const double y = y3[16 * q * TL_WB + 4 * t];
				acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(y, W2t[m][4 * t + g], acc[q], 0, 0, 0);
			}
		}
	}
#pragma unroll
	for(int q = 0; q < TL_WRB; q++)
#pragma unroll
		for(int e = 0; e < 4; e++){
			const unsigned r = rw0 + 16 * q + 4 * e + g;
			if(r < n && col < n) Z[(size_t)col * n + r] = acc[q][e];
		}
}

// ---- order, Boltzmann factors, partition sum -------------------------------------------------
template <int BT>
__global__ __launch_bounds__(BT) void tl_sort_kernel(const double *__restrict__ evall, unsigned n, double *__restrict__ evs_all,
		double *__restrict__ ems_all, int *__restrict__ perm_all, double *__restrict__ res)
{
	extern __shared__ double sh[];
	double *ev = sh, *red = sh + n;
	const unsigned s = blockIdx.x;
	for(unsigned i = threadIdx.x; i < n; i += BT) ev[i] = evall[(size_t)s * n + i];
	__syncthreads();
	for(unsigned i = threadIdx.x; i < n; i += BT){
		const double ei = ev[i];
		unsigned rank = 0;
		for(unsigned j = 0; j < n; j++){
			const double ej = ev[j];
			rank += (ej < ei || (ej == ei && j < i)) ? 1u : 0u;
		}
		evs_all[(size_t)s * n + rank] = ei;
		ems_all[(size_t)s * n + rank] = exp(-ei);
		perm_all[(size_t)s * n + rank] = (int)i;
	}
	__syncthreads();
	double z = 0;
	for(unsigned i = threadIdx.x; i < n; i += BT) z += ems_all[(size_t)s * n + i];
	z = block_sum<BT>(z, red);
	if(threadIdx.x == 0) res[s] = z / n;
}

// ---- sorted eigenvectors and W_c = G_c V ---------------------------------------------------
// (I_a v)[r] = f_a(r) v[nb(r,2a)] - f_a(nb(r,2a+1)) v[nb(r,2a+1)],  f_a(i) = t_a - kappa_a x[3i+a]
__global__ __launch_bounds__(256) void tl_current_kernel(const double *__restrict__ Zall, const int *__restrict__ perm_all,
		const double *__restrict__ xs, const int *__restrict__ nb, unsigned n, double t0, double t1, double t2,
		double k0, double k1, double k2, double c00, double c01, double c02, double c10, double c11, double c12,
		double *__restrict__ Vall, double *__restrict__ W0all, double *__restrict__ W1all)
{
	extern __shared__ double v[];
	const unsigned s = blockIdx.y, jj = blockIdx.x;
	const int src = perm_all[(size_t)s * n + jj];
	const double *z = Zall + (size_t)s * n * n + (size_t)src * n;
	const double *x = xs + (size_t)s * 3 * n;
	for(unsigned r = threadIdx.x; r < n; r += blockDim.x){
		const double a = z[r];
		v[r] = a;
		Vall[(size_t)s * n * n + (size_t)jj * n + r] = a;
	}
	__syncthreads();
	const double tk[3] = {t0, t1, t2}, kk[3] = {k0, k1, k2};
	const double C0[3] = {c00, c01, c02}, C1[3] = {c10, c11, c12};
	for(unsigned r = threadIdx.x; r < n; r += blockDim.x){
		double w0 = 0, w1 = 0;
#pragma unroll
		for(int a = 0; a < 3; a++){
			const unsigned fw = (unsigned)nb[(2 * a) * n + r], bw = (unsigned)nb[(2 * a + 1) * n + r];
			const double cur = (tk[a] - kk[a] * x[(size_t)r * 3 + a]) * v[fw] - (tk[a] - kk[a] * x[(size_t)bw * 3 + a]) * v[bw];
			w0 += C0[a] * cur;
			w1 += C1[a] * cur;
		}
		W0all[(size_t)s * n * n + (size_t)jj * n + r] = w0;
		W1all[(size_t)s * n * n + (size_t)jj * n + r] = w1;
	}
}

// ---- pair tiles: J_c(j, i) = v_j . W_c[:, i], binned --------------------------------------
// grid (i-block, j-block, sample), 64 x 64 tiles, 4 waves.  The tile is a GEMM over the N rows:
// wave w computes the 32 x 32 quarter (j: 32 (w / 2), i: 32 (w % 2)) of J_0 and J_1 with
// v_mfma_f64_16x16x4_f64, 2 x 2 blocks of 16 x 16, from 16-row panels of V, W_0, W_1 staged in
// LDS.  Operand layout of the instruction: A[m][k] in lane m + 16 k, B[k][n] in lane n + 16 k,
// D[m][n] in lane n + 16 (m % 4), element m / 4.
typedef double tl_v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void tl_pairs_kernel(const double *__restrict__ Vall, const double *__restrict__ W0all,
		const double *__restrict__ W1all, const double *__restrict__ evs_all, const double *__restrict__ ems_all,
		unsigned n, double width, unsigned nbins, double *__restrict__ part)
{
	__shared__ double Vs[TL_KC][TL_TILE + 1], W0s[TL_KC][TL_TILE + 1], W1s[TL_KC][TL_TILE + 1];
	__shared__ unsigned ebin[TL_TILE * TL_TILE];
	__shared__ double eval[TL_TILE * TL_TILE];
	__shared__ unsigned brange[2];
	__shared__ uint2 rrange[TL_TILE];
	__shared__ double hacc[TL_HB];
	const unsigned ib = blockIdx.x, jb = blockIdx.y, s = blockIdx.z, nblk = gridDim.x;
	const unsigned tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const unsigned i0 = ib * TL_TILE, j0 = jb * TL_TILE;
	const double *evs = evs_all + (size_t)s * n, *ems = ems_all + (size_t)s * n;
	double *out = part + (((size_t)s * nblk + jb) * nblk + ib) * nbins;
	// sorted order: E_j >= E_i needs j >= i except for equal eigenvalues across the blocks
	const unsigned jlast = min(j0 + TL_TILE, n) - 1;
	if(jb < ib && !(evs[jlast] >= evs[i0])){
		for(unsigned b = tid; b < nbins; b += 256) out[b] = 0;
		return;
	}
	const double *V = Vall + (size_t)s * n * n, *W0 = W0all + (size_t)s * n * n, *W1 = W1all + (size_t)s * n * n;
	const unsigned wj = 32 * (wave >> 1), wi = 32 * (wave & 1);      // this wave's quarter
	const unsigned m16 = lane & 15, k4 = lane >> 4;
	tl_v4d acc0[2][2], acc1[2][2];
#pragma unroll
	for(int a = 0; a < 2; a++)
#pragma unroll
		for(int b = 0; b < 2; b++){
			acc0[a][b] = tl_v4d{0, 0, 0, 0};
			acc1[a][b] = tl_v4d{0, 0, 0, 0};
		}
	if(tid == 0){
		brange[0] = TL_NONE;
		brange[1] = 0;
	}
	// the next stage's rows are loaded into registers while this stage's products run
	double pv[4], pw0[4], pw1[4];
	auto fetch = [&](unsigned r0){
#pragma unroll
		for(int q = 0; q < 4; q++){
			const unsigned e = tid + q * 256, cc = e / TL_KC, rr = e % TL_KC;
			const unsigned r = r0 + rr;
			const unsigned jc = j0 + cc, ic = i0 + cc;
			pv[q] = (jc < n && r < n) ? V[(size_t)jc * n + r] : 0.0;
			pw0[q] = (ic < n && r < n) ? W0[(size_t)ic * n + r] : 0.0;
			pw1[q] = (ic < n && r < n) ? W1[(size_t)ic * n + r] : 0.0;
		}
	};
	fetch(0);
	for(unsigned r0 = 0; r0 < n; r0 += TL_KC){
#pragma unroll
		for(int q = 0; q < 4; q++){
			const unsigned e = tid + q * 256, cc = e / TL_KC, rr = e % TL_KC;
			Vs[rr][cc] = pv[q];
			W0s[rr][cc] = pw0[q];
			W1s[rr][cc] = pw1[q];
		}
		__syncthreads();
		if(r0 + TL_KC < n) fetch(r0 + TL_KC);
#pragma unroll
		for(int kq = 0; kq < TL_KC / 4; kq++){
			const unsigned kk = 4 * kq + k4;
			double av[2], b0[2], b1[2];
#pragma unroll
			for(int a = 0; a < 2; a++){
				av[a] = Vs[kk][wj + 16 * a + m16];
				b0[a] = W0s[kk][wi + 16 * a + m16];
				b1[a] = W1s[kk][wi + 16 * a + m16];
			}
#pragma unroll
			for(int a = 0; a < 2; a++)
#pragma unroll
				for(int b = 0; b < 2; b++){
					acc0[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], b0[b], acc0[a][b], 0, 0, 0);
					acc1[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], b1[b], acc1[a][b], 0, 0, 0);
				}
		}
		__syncthreads();
	}
	const double wn = width * n;
	unsigned bmin = TL_NONE, bmax = 0;
#pragma unroll
	for(int a = 0; a < 2; a++)
#pragma unroll
		for(int b = 0; b < 2; b++)
#pragma unroll
			for(int e = 0; e < 4; e++){
				const unsigned jl = wj + 16 * a + 4 * e + k4, il = wi + 16 * b + m16;   // tile-local (j, i)
				const unsigned j = j0 + jl, i = i0 + il;
				unsigned bin = TL_NONE;
				double val = 0;
				if(j < n && i < n && evs[j] >= evs[i]){
					const double fr = (evs[j] - evs[i]) / width + 0.5;
					if(fr < (double)nbins){
						bin = (unsigned)fr;
						const double x0 = acc0[a][b][e], x1 = acc1[a][b][e];
						const double inm = x0 * x0 + x1 * x1;
						val = (ems[i] + ems[j]) * inm / wn;
						bmin = min(bmin, bin);
						bmax = max(bmax, bin);
					}
				}
				ebin[jl * TL_TILE + il] = bin;
				eval[jl * TL_TILE + il] = val;
			}
	if(bmin != TL_NONE){
		atomicMin(&brange[0], bmin);
		atomicMax(&brange[1], bmax);
	}
	__syncthreads();
	// Along a row (fixed j) the eigenvalues E_i ascend, so the bins are TL_NONE (difference past
	// the last bin), then non-increasing, then TL_NONE (E_i > E_j or i >= n): each bin is one run
	// per row.  Row ranges first, then the runs are added to their bins row by row.
	if(tid < TL_TILE){
		const unsigned *rb = ebin + tid * TL_TILE;
		unsigned lo = TL_TILE, hi = 0;
		for(unsigned il = 0; il < TL_TILE; il++)
			if(rb[il] != TL_NONE){
				lo = min(lo, il);
				hi = il + 1;
			}
		rrange[tid] = make_uint2(lo, hi);
	}
	__syncthreads();
	const unsigned blo = brange[0], bhi = brange[1];
	if(blo != TL_NONE && bhi - blo < TL_HB){
		// A row's runs fall in distinct bins, so one wave adds the runs of row after row (the LDS
		// operations of a wave complete in order): every bin adds its rows' run sums in row order,
		// each run summed with i ascending.
		for(unsigned b = tid; b < TL_HB; b += 256) hacc[b] = 0;
		__syncthreads();
		if(wave == 0)
			for(unsigned jl = 0; jl < TL_TILE; jl++){
				const uint2 rr = rrange[jl];
				const unsigned *rb = ebin + jl * TL_TILE;
				const unsigned il = lane;
				if(il >= rr.x && il < rr.y){
					const unsigned bin = rb[il];
					if(bin != TL_NONE && (il == rr.x || rb[il - 1] != bin)){
						double sum = 0;
						for(unsigned q = il; q < rr.y && rb[q] == bin; q++) sum += eval[jl * TL_TILE + q];
						hacc[bin - blo] += sum;
					}
				}
			}
		__syncthreads();
		for(unsigned b = tid; b < nbins; b += 256) out[b] = (b >= blo && b <= bhi) ? hacc[b - blo] : 0.0;
		return;
	}
	// wide tiles: every bin finds its run in each row by binary search and adds it row by row
	for(unsigned b = tid; b < nbins; b += 256){
		double sum = 0;
		if(blo != TL_NONE && b >= blo && b <= bhi)
			for(unsigned jl = 0; jl < TL_TILE; jl++){
				const uint2 rr = rrange[jl];
				if(rr.x >= rr.y) continue;
				const unsigned *rb = ebin + jl * TL_TILE;
				if(b > rb[rr.x] || b < rb[rr.y - 1]) continue;
				// first il with bin <= b, then first il with bin < b
				unsigned p = rr.x, q = rr.y;
				while(p < q){
					const unsigned mid = (p + q) >> 1;
					if(rb[mid] <= b) q = mid;
					else p = mid + 1;
				}
				unsigned t = p, u = rr.y;
				while(t < u){
					const unsigned mid = (t + u) >> 1;
					if(rb[mid] < b) u = mid;
					else t = mid + 1;
				}
				for(unsigned il = p; il < t; il++) sum += eval[jl * TL_TILE + il];
			}
		out[b] = sum;
	}
}

// Fixed-order sum of the tile histograms in two stages: chunk c of TL_RC tiles is summed by one
// workgroup (wave w takes tiles w, w+4, ... of the chunk for 64 bins, the four wave sums are
// added in wave order), then every bin adds its chunk sums in chunk order.
constexpr unsigned TL_RC = 512;

__global__ __launch_bounds__(256) void tl_reduce_chunks_kernel(const double *__restrict__ part, unsigned nparts, unsigned nbins,
		double *__restrict__ red)
{
	__shared__ double ws[4][64];
	const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const unsigned b = blockIdx.x * 64 + lane, c = blockIdx.y;
	const unsigned t0 = c * TL_RC, t1 = min(t0 + TL_RC, nparts);
	double s = 0;
	if(b < nbins)
		for(unsigned t = t0 + w; t < t1; t += 4) s += part[(size_t)t * nbins + b];
	ws[w][lane] = s;
	__syncthreads();
	if(w == 0 && b < nbins) red[(size_t)c * nbins + b] = ((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane];
}

__global__ void tl_reduce_kernel(const double *__restrict__ red, unsigned nchunks, unsigned nbins, double *__restrict__ acc)
{
	const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
	if(b >= nbins) return;
	double s = 0;
	for(unsigned c = 0; c < nchunks; c++) s += red[(size_t)c * nbins + b];
	acc[b] += s;
}

__global__ void tl_zero_kernel(double *a, unsigned n)
{
	const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n) a[i] = 0;
}

// ---------------------------------------------------------------------------------------------
static int tl_reserve(ephmc_ctx *c, unsigned B, unsigned nbins)
{
	TlWork &w = c->tl;
	const unsigned n = c->ns, nblk = (n + TL_TILE - 1) / TL_TILE;
	if(w.cap_b >= B && w.cap_nbins >= nbins) return EPHMC_OK;
	tl_release(c);
	const size_t mat = (size_t)n * n;
	w.rot_cap = std::max<size_t>(3 * mat / 2, 64);
	w.sw_cap = 16 * n + 64;
	auto al = [&](void **p, size_t bytes) -> int {
		if(hipMalloc(p, bytes) != hipSuccess){
			(void)hipGetLastError();
			set_error("transient-localisation workspace: hipMalloc(%zu) failed", bytes);
			return EPHMC_ERR_ALLOC;
		}
		return EPHMC_OK;
	};
	int rc;
	if((rc = al((void **)&w.A, B * mat * 8)) || (rc = al((void **)&w.Z, B * mat * 8))
		|| (rc = al((void **)&w.rot, B * w.rot_cap * sizeof(double2))) || (rc = al((void **)&w.sw, B * (size_t)w.sw_cap * sizeof(int4)))
		|| (rc = al((void **)&w.nsw, B * sizeof(unsigned))) || (rc = al((void **)&w.status, B * sizeof(int)))
		|| (rc = al((void **)&w.x, B * 3 * (size_t)n * 8)) || (rc = al((void **)&w.d, B * (size_t)n * 8))
		|| (rc = al((void **)&w.e, B * (size_t)n * 8)) || (rc = al((void **)&w.tau, B * (size_t)n * 8))
		|| (rc = al((void **)&w.ev, B * (size_t)n * 8)) || (rc = al((void **)&w.evs, B * (size_t)n * 8))
		|| (rc = al((void **)&w.ems, B * (size_t)n * 8)) || (rc = al((void **)&w.perm, B * (size_t)n * sizeof(int)))
		|| (rc = al((void **)&w.res, B * 8)) || (rc = al((void **)&w.part, B * (size_t)nblk * nblk * nbins * 8))
		|| (rc = al((void **)&w.acc, (size_t)nbins * 8))
		|| (rc = al((void **)&w.red, ((B * (size_t)nblk * nblk + TL_RC - 1) / TL_RC) * nbins * 8))){
		tl_release(c);
		return rc;
	}
	// A is free once the back-transformation has run, the rotations once they are applied
	// (2 N^2 doubles fit in 1.5 N^2 double pairs)
	w.V = w.A;
	w.W0 = (double *)w.rot;
	w.W1 = (double *)w.rot + B * mat;
	w.cap_b = B;
	w.cap_nbins = nbins;
	return EPHMC_OK;
}

void tl_release(ephmc_ctx *c)
{
	TlWork &w = c->tl;
	void *bufs[] = {w.A, w.Z, w.rot, w.red, w.sw, w.nsw, w.status, w.x, w.d, w.e, w.tau, w.ev, w.evs, w.ems, w.perm,
		w.res, w.part, w.acc};
	for(void *b : bufs) if(b) (void)hipFree(b);
	w = TlWork();
}

} // namespace ephmc

using namespace ephmc;

extern "C" int ephmc_tl_hist(ephmc_ctx *c, const double *normals, unsigned nsamples, const double *contr_mat, double width,
		unsigned nbins, double *greens, double *results)
{
	if(!c || !normals || !contr_mat || !greens || !results){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(nsamples == 0) return EPHMC_OK;
	if(nbins == 0 || nbins > 65536){ set_error("nbins must be in [1, 65536], got %u", nbins); return EPHMC_ERR_ARG; }
	if(!(width > 0)){ set_error("bin width must be positive, got %g", width); return EPHMC_ERR_ARG; }
	const unsigned n = c->ns;
	if(n > 2 * TL_BT){ set_error("transient localisation supports N <= %d sites", 2 * TL_BT); return EPHMC_ERR_ARG; }
	const size_t mat = (size_t)n * n;
	const unsigned nblk = (n + TL_TILE - 1) / TL_TILE;
	// samples per batch: about 100 GB of workspace (5.5 N^2 doubles per sample; EPHMC_TL_BATCH_MB
	// sets another budget in MiB), at most 2048, so that the one-wave-per-sample QL recurrence has
	// two waves on every SIMD
	size_t mb = (size_t)100 << 10;
	if(const char *v = getenv("EPHMC_TL_BATCH_MB")) mb = std::max<size_t>(1, strtoull(v, nullptr, 10));
	const size_t per = (size_t)(5.5 * mat * 8) + (size_t)nblk * nblk * nbins * 8;
	unsigned B = (unsigned)std::max<size_t>(1, (mb << 20) / per);
	B = std::min(B, 2048u);
	B = std::min(B, nsamples);
	EPHMC_TRY(tl_reserve(c, B, nbins));
	TlWork &w = c->tl;
	EPHMC_TRY(timer_begin(c, "tl"));
	tl_zero_kernel<<<(nbins + 255) / 256, 256, 0, c->stream>>>(w.acc, nbins);
	EPHMC_LAUNCHED();
	std::vector<int> status(B);
	for(unsigned s0 = 0; s0 < nsamples; s0 += B){
		const unsigned nb_ = std::min(B, nsamples - s0);
		EPHMC_HIP(hipMemcpyAsync(w.x, normals + (size_t)s0 * 3 * n, (size_t)nb_ * 3 * n * 8, hipMemcpyHostToDevice, c->stream));
		{
			const dim3 grid((unsigned)std::min<size_t>((mat + 255) / 256, 1024), nb_);
			tl_build_kernel<<<grid, 256, 0, c->stream>>>(w.x, c->d_nb, c->d_link, c->cfg.t[0], c->cfg.t[1], c->cfg.t[2],
				c->cfg.kappa[0], c->cfg.kappa[1], c->cfg.kappa[2], -c->cfg.mu, n, w.A);
			EPHMC_LAUNCHED();
		}
		EPHMC_TRY(timer_begin(c, "tl_tridiag"));
		{
			if(n <= TL_BT){
				const size_t lds = tl_panel_lds_doubles<TL_BT, TL_PNB>(n) * 8;
				(void)hipFuncSetAttribute((const void *)tl_tridiag_panel_kernel<TL_BT, 1, TL_PNB>,
					hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
				tl_tridiag_panel_kernel<TL_BT, 1, TL_PNB><<<nb_, TL_BT, lds, c->stream>>>(w.A, n, w.d, w.e, w.tau);
			}else{
				const size_t lds = (3 * (size_t)n + TL_BT / 64) * 8;
				if(n <= TL_BT) tl_tridiag_kernel<TL_BT, 1><<<nb_, TL_BT, lds, c->stream>>>(w.A, n, w.d, w.e, w.tau);
				else tl_tridiag_kernel<TL_BT, 2><<<nb_, TL_BT, lds, c->stream>>>(w.A, n, w.d, w.e, w.tau);
			}
			EPHMC_LAUNCHED();
		}
		EPHMC_TRY(timer_end(c, "tl_tridiag"));
		EPHMC_TRY(timer_begin(c, "tl_ql"));
		tl_ql_scalar_kernel<<<nb_, 64, 2 * (size_t)n * 8, c->stream>>>(w.d, w.e, n, w.ev, w.rot, w.rot_cap, w.sw, w.sw_cap, w.nsw,
			w.status);
		EPHMC_LAUNCHED();
		tl_ql_apply_kernel<TL_G, TL_PF, TL_QC><<<dim3((n + 255) / 256, nb_), 256, 0, c->stream>>>(w.Z, n, w.rot, w.rot_cap, w.sw, w.sw_cap, w.nsw);
		EPHMC_LAUNCHED();
		EPHMC_TRY(timer_end(c, "tl_ql"));
		EPHMC_TRY(timer_begin(c, "tl_back"));
		{
			// compact-WY blocks on the matrix cores (N <= 1024), the reflector copy and the T factors in
			// the rotation buffer, which is free once the rotations are applied; else one reflector at a time
			const unsigned nblkw = n > 2 ? (n - 2 + TL_WB - 1) / TL_WB : 0;
			const size_t ybsz = (size_t)nblkw * ((n + 15) & ~15u) * TL_WB;
			double *yball = (double *)(void *)w.rot, *tall = yball + (size_t)nb_ * ybsz;
			const bool wy = n >= 3 && n <= TL_WW * 16 * TL_WRB && ybsz + (size_t)nblkw * TL_WB * TL_WB <= 2 * w.rot_cap;
			if(wy){
				tl_yb_kernel<<<dim3((n + 31) / 32, nblkw, nb_), 256, 0, c->stream>>>(w.A, n, nblkw, yball);
				EPHMC_LAUNCHED();
				tl_wy_t_kernel<<<dim3(nblkw, nb_), 256, 0, c->stream>>>(w.A, w.tau, n, tall);
				EPHMC_LAUNCHED();
				tl_wy_apply_kernel<<<dim3((n + TL_WC - 1) / TL_WC, nb_), 64 * TL_WW, 0, c->stream>>>(yball, tall, n, nblkw, w.Z);
			}else if(n <= 1024) tl_back_kernel<16, 4><<<dim3((n + 15) / 16, nb_), 256, 0, c->stream>>>(w.A, w.tau, w.Z, n);
			else tl_back_kernel<32, 2><<<dim3((n + 7) / 8, nb_), 256, 0, c->stream>>>(w.A, w.tau, w.Z, n);
			EPHMC_LAUNCHED();
		}
		EPHMC_TRY(timer_end(c, "tl_back"));
		tl_sort_kernel<TL_BT><<<nb_, TL_BT, ((size_t)n + TL_BT / 64) * 8, c->stream>>>(w.ev, n, w.evs, w.ems, w.perm, w.res);
		EPHMC_LAUNCHED();
		tl_current_kernel<<<dim3(n, nb_), 256, (size_t)n * 8, c->stream>>>(w.Z, w.perm, w.x, c->d_nb, n, c->cfg.t[0], c->cfg.t[1],
			c->cfg.t[2], c->cfg.kappa[0], c->cfg.kappa[1], c->cfg.kappa[2], contr_mat[0], contr_mat[1], contr_mat[2],
			contr_mat[3], contr_mat[4], contr_mat[5], w.V, w.W0, w.W1);
		EPHMC_LAUNCHED();
		EPHMC_TRY(timer_begin(c, "tl_pairs"));
		tl_pairs_kernel<<<dim3(nblk, nblk, nb_), 256, 0, c->stream>>>(w.V, w.W0, w.W1, w.evs, w.ems, n, width, nbins, w.part);
		EPHMC_LAUNCHED();
		{
			const unsigned nparts = nb_ * nblk * nblk, nch = (nparts + TL_RC - 1) / TL_RC;
			tl_reduce_chunks_kernel<<<dim3((nbins + 63) / 64, nch), 256, 0, c->stream>>>(w.part, nparts, nbins, w.red);
			EPHMC_LAUNCHED();
			tl_reduce_kernel<<<(nbins + 255) / 256, 256, 0, c->stream>>>(w.red, nch, nbins, w.acc);
			EPHMC_LAUNCHED();
		}
		EPHMC_TRY(timer_end(c, "tl_pairs"));
		EPHMC_HIP(hipMemcpyAsync(results + s0, w.res, (size_t)nb_ * 8, hipMemcpyDeviceToHost, c->stream));
		EPHMC_HIP(hipMemcpyAsync(status.data(), w.status, (size_t)nb_ * sizeof(int), hipMemcpyDeviceToHost, c->stream));
		EPHMC_HIP(hipStreamSynchronize(c->stream));
		for(unsigned k = 0; k < nb_; k++)
			if(status[k]){
				set_error("eigensolver of disorder sample %u: %s", s0 + k,
					status[k] == 1 ? "QL iteration did not converge in 60 sweeps" : "rotation record capacity exceeded");
				return EPHMC_ERR_NUMERIC;
			}
	}
	EPHMC_TRY(timer_end(c, "tl"));
	std::vector<double> acc(nbins);
	EPHMC_HIP(hipMemcpyAsync(acc.data(), w.acc, (size_t)nbins * 8, hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	for(unsigned b = 0; b < nbins; b++) greens[b] += acc[b];
	return EPHMC_OK;
}
