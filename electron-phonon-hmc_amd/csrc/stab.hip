// stab.hip -- the slice product with periodic FP64 stabilisation: log det and inverse of
// 1 + B_0 B_1 ... B_{nt-1} without forming the product's full dynamic range in one matrix.
//
// The product is kept as U diag(d) V (U orthogonal, d > 0 the column scales, V of O(1) entries) and
// grown by blocks of k slices, in the reference's order (slices nt-1 first, applied from the left,
// construct_M_hat organic_dense_matrix.c:256-258):
//     Y = B_lo ... B_hi U                  (the chain kernels, in the slice precision)
//     Z = Y diag(d) Pi                     (columns sorted by norm, largest first: pre-pivoting)
//     Z = Q R                              (FP64 Householder QR, one workgroup)
//     U <- Q,  d <- |diag R|,  V <- (diag(d)^{-1} R) Pi^T V
// Every block's dynamic range is that of k slices only, so an FP32 block loses nothing the FP64
// factorisation cannot hold.  At the end, with D_b = max(d, 1) and D_s = min(d, 1),
//     1 + U D V = U D_b X,   X = D_b^{-1} U^T + D_s V,
// so log|det(1 + U D V)| = sum log D_b + log|det X| (det U = +-1) and
//     (1 + U D V)^{-1} = X^{-1} D_b^{-1} U^T;
// X is well conditioned however wide d is, and its inverse is the Gauss-Jordan kernel's.
#include "internal.h"

namespace ephmc {

constexpr int QR_BT = 1024;

// In-place pre-pivoted Householder QR of Z = Y diag(d) Pi (one workgroup; N <= 2048).
// Y: column-major N x N in R (the block product); d: column scales.  Out: zr holds R on and above
// the diagonal and the Householder vectors (unit leading entry implied) below it; tau[j]; perm[k]
// = source column of column k; nref = number of nonzero reflectors (det Q = (-1)^nref).
template <typename R>
__global__ __launch_bounds__(QR_BT) void qr_prepivot_kernel(const R *__restrict__ Y, const double *__restrict__ d,
		unsigned n, double *__restrict__ zr, double *__restrict__ tau, int *__restrict__ perm, int *__restrict__ nref)
{
	extern __shared__ double qsh[];
	double *key = qsh;                              // n: log of the scaled column norms
	__shared__ double nsig[2], nalpha[2];           // sigma and alpha of the column of the next step
	const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = QR_BT / 64;
	// column keys: log(d_c) + log|Y_c| (logs: the scales may span far more than an FP64 square)
	for(unsigned c = tid; c < n; c += QR_BT){
		double s = 0;
		const R *y = Y + (size_t)c * n;
		for(unsigned r = 0; r < n; r++){ const double v = (double)y[r]; s = fma(v, v, s); }
		key[c] = s > 0 ? log(d[c]) + 0.5 * log(s) : -HUGE_VAL;
	}
	__syncthreads();
	// rank by key, largest first, ties by column index
	for(unsigned c = tid; c < n; c += QR_BT){
		const double kc = key[c];
		unsigned rank = 0;
		for(unsigned o = 0; o < n; o++){
			const double ko = key[o];
			rank += (ko > kc || (ko == kc && o < c)) ? 1u : 0u;
		}
		perm[rank] = (int)c;
	}
	__syncthreads();
	// Z = Y diag(d) Pi
	for(size_t e = tid; e < (size_t)n * n; e += QR_BT){
		const unsigned k = (unsigned)(e / n), r = (unsigned)(e - (size_t)k * n);
		const int src = perm[k];
		zr[e] = d[src] * (double)Y[(size_t)src * n + r];
	}
	__syncthreads();
	// Each step j needs column j below the diagonal (staged in LDS), its squared norm sigma and its
	// diagonal entry alpha; the wave that updates column j + 1 in step j stages them for step j + 1, so
	// a step is one pass over the trailing columns and one barrier.  Column 0 first:
	double *const vb0 = qsh + n;                    // two n-double buffers at qsh + n + b n
	if(wave == 0){
		double s = 0;
		for(unsigned r = 1 + lane; r < n; r += 64){
			const double v = zr[r];
			vb0[r] = v;
			s = fma(v, v, s);
		}
#pragma unroll
		for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
		if(lane == 0){
			nsig[0] = s;
			nalpha[0] = zr[0];
		}
	}
	__syncthreads();
	int refl = 0;
	for(unsigned j = 0; j < n; j++){
		const int b = j & 1;
		const double *vc = vb0 + (size_t)b * n;
		double *vn = vb0 + (size_t)(b ^ 1) * n;
		double *zj = zr + (size_t)j * n;
		// H_j = 1 - t v v^T with v_j = 1, v_r = z_rj / (alpha - beta) (the same values in every thread)
		const double sig = nsig[b], alpha = nalpha[b];
		double t = 0, scale = 0, beta = alpha;
		if(sig > 0){
			const double nrm = sqrt(fma(alpha, alpha, sig));
			beta = alpha >= 0 ? -nrm : nrm;
			t = (beta - alpha) / beta;
			scale = 1.0 / (alpha - beta);
		}
		if(tid == 0){
			tau[j] = t;
			zj[j] = beta;
		}
		if(t != 0){
			refl++;
			for(unsigned r = j + 1 + tid; r < n; r += QR_BT) zj[r] = vc[r] * scale;     // the stored v
			// columns c > j: w = z_jc + sum_{r > j} v_r z_rc;  z_.c -= t w v   (one wave per column;
			// wave 0 takes column j + 1 first and stages it for step j + 1)
			for(unsigned c = j + 1 + wave; c < n; c += nw){
				double *zc = zr + (size_t)c * n;
				double w = 0;
				for(unsigned r = j + 1 + lane; r < n; r += 64) w = fma(vc[r] * scale, zc[r], w);
#pragma unroll
				for(int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
				w += zc[j];
				const double tw = t * w;
				if(lane == 0) zc[j] -= tw;
				if(c == j + 1){
					double s = 0, diag = 0;
					for(unsigned r = j + 1 + lane; r < n; r += 64){
						const double nv = fma(-tw, vc[r] * scale, zc[r]);
						zc[r] = nv;
						if(r == j + 1) diag = nv;
						else{
							vn[r] = nv;
							s = fma(nv, nv, s);
						}
					}
#pragma unroll
					for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
					if(lane == 0){
						nsig[b ^ 1] = s;
						nalpha[b ^ 1] = diag;
					}
				}else{
					for(unsigned r = j + 1 + lane; r < n; r += 64) zc[r] = fma(-tw, vc[r] * scale, zc[r]);
				}
			}
		}else if(wave == 0 && j + 1 < n){
			// no reflector: column j + 1 is staged as it stands
			const double *zc = zr + (size_t)(j + 1) * n;
			double s = 0;
			for(unsigned r = j + 2 + lane; r < n; r += 64){
				const double v = zc[r];
				vn[r] = v;
				s = fma(v, v, s);
			}
#pragma unroll
			for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
			if(lane == 0){
				nsig[b ^ 1] = s;
				nalpha[b ^ 1] = zc[j + 1];
			}
		}
		__syncthreads();
	}
	if(tid == 0) *nref = refl;
}

// Q = H_0 H_1 ... H_{n-1} (the reflectors stored in zr, tau) into q, column-major; one workgroup.
// Reflector j is read from LDS; reflector j - 1 is loaded while j is applied and staged after it
// (two n-double buffers), so a step is one pass over the columns and one barrier.
__global__ __launch_bounds__(QR_BT) void qr_form_q_kernel(const double *__restrict__ zr, const double *__restrict__ tau,
		unsigned n, double *__restrict__ q)
{
	extern __shared__ double vsh[];                 // 2 n: v_r (r > j) of reflector j at vsh + (j & 1) n
	constexpr unsigned PR = 2048 / QR_BT;           // rows of the next reflector per thread (n <= 2048)
	const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = QR_BT / 64;
	for(size_t e = tid; e < (size_t)n * n; e += QR_BT){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		q[e] = r == c ? 1.0 : 0.0;
	}
	// reflector n - 1 has no rows below its diagonal
	__syncthreads();
	for(int j = (int)n - 1; j >= 0; j--){
		const unsigned b = (unsigned)j & 1u;
		const double *vc = vsh + (size_t)b * n;
		double *vn = vsh + (size_t)(b ^ 1u) * n;
		// the next reflector's rows r > j - 1, in flight during this step's columns
		double pre[PR];
#pragma unroll
		for(unsigned k = 0; k < PR; k++){
			const unsigned r = (unsigned)j + tid + QR_BT * k;
			pre[k] = (j > 0 && r < n) ? zr[(size_t)(j - 1) * n + r] : 0.0;
		}
		const double t = tau[j];
		if(t != 0){
			for(unsigned c = (unsigned)j + wave; c < n; c += nw){
				double *qc = q + (size_t)c * n;
				double w = 0;
				for(unsigned r = j + 1 + lane; r < n; r += 64) w = fma(vc[r], qc[r], w);
#pragma unroll
				for(int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
				w += qc[j];
				const double tw = t * w;
				if(lane == 0) qc[j] -= tw;
				for(unsigned r = j + 1 + lane; r < n; r += 64) qc[r] = fma(-tw, vc[r], qc[r]);
			}
		}
#pragma unroll
		for(unsigned k = 0; k < PR; k++){
			const unsigned r = (unsigned)j + tid + QR_BT * k;
			if(j > 0 && r < n) vn[r] = pre[k];
		}
		__syncthreads();
	}
}

// d_new[k] = |R_kk|; V <- (diag(d_new)^{-1} R) Pi^T V, one workgroup per column of V (in place:
// the column is staged, permuted, in LDS first)
__global__ __launch_bounds__(256) void qr_update_v_kernel(const double *__restrict__ zr, const int *__restrict__ perm,
		unsigned n, double *__restrict__ v, double *__restrict__ d_new)
{
	extern __shared__ double col[];
	const unsigned c = blockIdx.x;
	double *vc = v + (size_t)c * n;
	for(unsigned k = threadIdx.x; k < n; k += blockDim.x) col[k] = vc[perm[k]];
	__syncthreads();
	for(unsigned i = threadIdx.x; i < n; i += blockDim.x){
		const double rii = zr[(size_t)i * n + i];
		double s = 0;
		for(unsigned k = i; k < n; k++) s = fma(zr[(size_t)k * n + i], col[k], s);
		vc[i] = s / fabs(rii);
		if(c == 0) d_new[i] = fabs(rii);
	}
}

// X = D_b^{-1} U^T + D_s V (column-major), D_b = max(d, 1), D_s = min(d, 1)
__global__ __launch_bounds__(256) void stab_x_kernel(const double *__restrict__ u, const double *__restrict__ v,
		const double *__restrict__ d, unsigned n, double *__restrict__ x)
{
	__shared__ double tile[32][33];
	const unsigned bx = blockIdx.x * 32, by = blockIdx.y * 32;
	const unsigned tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	// tile of U: rows bx.., columns by..  ->  transposed into X rows by.., columns bx..
	for(unsigned r = ty; r < 32; r += 8){
		const unsigned col = by + r, row = bx + tx;
		if(col < n && row < n) tile[r][tx] = u[(size_t)col * n + row];
	}
	__syncthreads();
	for(unsigned r = ty; r < 32; r += 8){
		const unsigned col = bx + r, row = by + tx;     // X(row, col) = U(col, row) / Db_row + Ds_row V(row, col)
		if(col < n && row < n){
			const double dr = d[row];
			const double db = dr > 1 ? dr : 1.0, ds = dr < 1 ? dr : 1.0;
			x[(size_t)col * n + row] = tile[tx][r] / db + ds * v[(size_t)col * n + row];
		}
	}
}

// out = A diag(1 / D_b) B^T (column-major N x N, D_b = max(d, 1)): 64 x 64 tiles, 256 threads of
// 4 x 4 outputs each, k in panels of 16 through LDS
__global__ __launch_bounds__(256) void gemm_abt_kernel(const double *__restrict__ a, const double *__restrict__ b,
		const double *__restrict__ d, unsigned n, double *__restrict__ out)
{
	__shared__ double as[16][65], bs[16][65];
	const unsigned r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
	const unsigned tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
	double acc[4][4] = {};
	for(unsigned k0 = 0; k0 < n; k0 += 16){
		for(unsigned e = threadIdx.x; e < 16 * 64; e += 256){
			const unsigned kk = e >> 6, i = e & 63;
			const unsigned k = k0 + kk;
			double sa = 0, sb = 0;
			if(k < n){
				const double dk = d[k] > 1 ? d[k] : 1.0;
				if(r0 + i < n) sa = a[(size_t)k * n + r0 + i] / dk;        // A(r0+i, k) / Db_k
				if(c0 + i < n) sb = b[(size_t)k * n + c0 + i];             // B(c0+i, k) = B^T(k, c0+i)
			}
			as[kk][i] = sa;
			bs[kk][i] = sb;
		}
		__syncthreads();
#pragma unroll
		for(int kk = 0; kk < 16; kk++){
			double av[4], bv[4];
#pragma unroll
			for(int q = 0; q < 4; q++){ av[q] = as[kk][tr + 16 * q]; bv[q] = bs[kk][tc + 16 * q]; }
#pragma unroll
			for(int p = 0; p < 4; p++)
#pragma unroll
				for(int q = 0; q < 4; q++) acc[p][q] = fma(av[p], bv[q], acc[p][q]);
		}
		__syncthreads();
	}
#pragma unroll
	for(int p = 0; p < 4; p++)
#pragma unroll
		for(int q = 0; q < 4; q++){
			const unsigned r = r0 + tr + 16 * p, cc = c0 + tc + 16 * q;
			if(r < n && cc < n) out[(size_t)cc * n + r] = acc[p][q];
		}
}

// log det(1 + U D V) = sum log D_b + log|det X| (X's in scal[SCAL_LOGDET] from the Gauss-Jordan);
// the sign picks up det U = (-1)^nref
__global__ __launch_bounds__(256) void stab_logdet_kernel(const double *__restrict__ d, unsigned n,
		const int *__restrict__ nref, double *__restrict__ scal)
{
	__shared__ double part[256];
	double s = 0;
	for(unsigned i = threadIdx.x; i < n; i += 256) s += d[i] > 1 ? log(d[i]) : 0.0;
	part[threadIdx.x] = s;
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1){
		if((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
		__syncthreads();
	}
	if(threadIdx.x == 0){
		scal[SCAL_LOGDET] += part[0];
		if(*nref & 1) scal[SCAL_SIGN] = -scal[SCAL_SIGN];
	}
}


// buffers of the stabilised product, allocated together on first use (or at creation when
// ephmc_config.stabilize is set)
int stab_reserve(ephmc_ctx *c)
{
	if(c->stab.U) return EPHMC_OK;
	const size_t mat = c->mat, n = c->ns;
	double **d[] = {&c->stab.U, &c->stab.V, &c->stab.ZR, &c->stab.X, &c->stab.d, &c->stab.tau};
	const size_t sz[] = {mat, mat, mat, mat, n, n};
	for(int i = 0; i < 6; i++){
		if(hipMalloc((void **)d[i], sz[i] * sizeof(double)) != hipSuccess){
			(void)hipGetLastError();
			stab_release(c);
			set_error("stabilised product: device buffers (%zu MB) could not be allocated", (4 * mat * 8) >> 20);
			return EPHMC_ERR_ALLOC;
		}
	}
	if(hipMalloc((void **)&c->stab.perm, n * sizeof(int)) != hipSuccess
		|| hipMalloc((void **)&c->stab.nref, sizeof(int)) != hipSuccess){
		(void)hipGetLastError();
		stab_release(c);
		set_error("stabilised product: device buffers could not be allocated");
		return EPHMC_ERR_ALLOC;
	}
	return EPHMC_OK;
}

void stab_release(ephmc_ctx *c)
{
	void *b[] = {c->stab.U, c->stab.V, c->stab.ZR, c->stab.X, c->stab.d, c->stab.tau, c->stab.perm, c->stab.nref};
	for(void *p : b) if(p) (void)hipFree(p);
	c->stab = StabWork{};
}

__global__ void fill_kernel(double *__restrict__ a, double v, size_t n)
{
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) a[e] = v;
}

// minv = (1 + B_0 ... B_{nt-1})^{-1} and its log det (scal[SCAL_LOGDET], sign in SCAL_SIGN) with the
// slice products in precision prec and an FP64 QR every k slices; the coefficients of the field
// must be current for prec.  Scratch: the buffers above, d_Tfin / d_Tfin32 (block products),
// d_gj (Gauss-Jordan).
int stable_logdet_inverse(ephmc_ctx *c, int prec, unsigned k, double *minv)
{
	const unsigned n = c->ns, nt = c->nt;
	if(n > 2048){ set_error("stabilised product: N = %u exceeds 2048", n); return EPHMC_ERR_UNSUPPORTED; }
	if(k == 0) k = 1;
	EPHMC_TRY(stab_reserve(c));
	StabWork &w = c->stab;
	EPHMC_TRY(set_identity(c, EPHMC_FP64, w.U));
	EPHMC_TRY(set_identity(c, EPHMC_FP64, w.V));
	fill_kernel<<<(n + 255) / 256, 256, 0, c->stream>>>(w.d, 1.0, n);
	EPHMC_LAUNCHED();
	void *Y = prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32;
	const size_t qlds = (size_t)n * sizeof(double);
	for(int hi = (int)nt - 1; hi >= 0; hi -= (int)k){
		const int cnt = std::min<int>((int)k, hi + 1);
		// Y = B_{hi-cnt+1} ... B_hi U: slices hi down to hi-cnt+1, each applied from the left
		EPHMC_TRY(chain_range(c, prec, w.U, Y, hi, -1, cnt));
		if(prec == EPHMC_FP64)
			qr_prepivot_kernel<double><<<1, QR_BT, 3 * qlds, c->stream>>>((const double *)Y, w.d, n, w.ZR, w.tau, w.perm, w.nref);
		else
			qr_prepivot_kernel<float><<<1, QR_BT, 3 * qlds, c->stream>>>((const float *)Y, w.d, n, w.ZR, w.tau, w.perm, w.nref);
		EPHMC_LAUNCHED();
		qr_form_q_kernel<<<1, QR_BT, 2 * qlds, c->stream>>>(w.ZR, w.tau, n, w.U);
		EPHMC_LAUNCHED();
		qr_update_v_kernel<<<n, 256, qlds, c->stream>>>(w.ZR, w.perm, n, w.V, w.d);
		EPHMC_LAUNCHED();
	}
	const dim3 tg((n + 31) / 32, (n + 31) / 32);
	stab_x_kernel<<<tg, 256, 0, c->stream>>>(w.U, w.V, w.d, n, w.X);
	EPHMC_LAUNCHED();
	EPHMC_TRY(gj_logdet_inverse(c, w.X));
	const dim3 gg((n + 63) / 64, (n + 63) / 64);
	gemm_abt_kernel<<<gg, 256, 0, c->stream>>>(w.X, w.U, w.d, n, minv);
	EPHMC_LAUNCHED();
	stab_logdet_kernel<<<1, 256, 0, c->stream>>>(w.d, n, w.nref, c->d_scal);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

} // namespace ephmc

using namespace ephmc;

extern "C" int ephmc_stable_logdet_inv(ephmc_ctx *c, const double *x, int precision, unsigned k, double *M_inv,
		double *logdet, int *sign)
{
	if(!c || !x || !M_inv){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(precision != EPHMC_FP64 && precision != EPHMC_MIXED){ set_error("precision"); return EPHMC_ERR_ARG; }
	if(c->cfg.zero_filling || c->cfg.single_particle){
		set_error("the stabilised product forms 1 + B..B: not for ZERO_FILLING / SINGLE_PARTICLE");
		return EPHMC_ERR_UNSUPPORTED;
	}
	if(precision == EPHMC_MIXED && !c->d_coef32){
		set_error("FP32 slices need a context created with EPHMC_MIXED");
		return EPHMC_ERR_ARG;
	}
	EPHMC_TRY(coefficients(c, x, precision == EPHMC_MIXED));
	EPHMC_TRY(stable_logdet_inverse(c, precision, k, M_inv));
	if(!logdet && !sign) return EPHMC_OK;
	double s[SCAL_COUNT];
	EPHMC_HIP(hipMemcpyAsync(s, c->d_scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	if(logdet) *logdet = s[SCAL_LOGDET];
	if(sign) *sign = s[SCAL_SIGN] < 0 ? -1 : 1;
	if(s[SCAL_SINGULAR] != 0 || s[SCAL_LOGDET] == -HUGE_VAL){
		set_error("singular matrix in the stabilised inverse; log|det| = -inf");
		return EPHMC_ERR_NUMERIC;
	}
	return EPHMC_OK;
}
