// schur.hip -- the SCHUR_FORM_INV branch of the log det and inverse (organic_dense_matrix.c:556-578,
// 599-603): B = M_hat - 1 in real Schur form B = Z T Z^T, the entries of T below a threshold set
// to zero when the spectrum is wide (max_ev eps / 1e-3 > 1) or has a real part below -1, and
//     log|det M_hat| = log|det(1 + T)|,   M_hat^{-1} = Z (1 + T)^{-1} Z^T.
//
// The Schur form is computed here (the reference calls LAPACK dgees): a Householder reduction to
// upper Hessenberg form, then Francis double-shift QR sweeps with the transformations accumulated
// into Z, in the bulge-chasing formulation of EISPACK's hqr2 (Wilkinson-Reinsch), with a 2 x 2
// block standardised by a rotation once it deflates (real pair: upper triangular; complex pair:
// equal diagonal entries), as dgees leaves them.  All of it runs in one 1024-thread workgroup:
// the control flow (deflation, shifts, reflectors) is decided by one thread and broadcast through
// LDS; every row, column and Schur-vector update is spread over the workgroup.  The Schur form is
// not unique (eigenvalue order, signs of Z's columns), so only what does not depend on those
// choices is compared with the plain inverse: log det and M_hat^{-1} where nothing is zeroed, and
// the zeroed result for spectra where the threshold separates the eigenvalues cleanly.
#include "internal.h"
#include <cfloat>

namespace ephmc {

constexpr int SC_BT = 1024;

__device__ __forceinline__ double hsign(double a, double b){ return b >= 0 ? fabs(a) : -fabs(a); }

// In place: B (column-major N x N, element (r, c) at B[c * n + r]) -> upper Hessenberg H with the
// Householder vectors of the reduction kept below the first subdiagonal (v_{j+1} = 1 implied);
// then Z = H_0 ... H_{n-3} and the entries below the subdiagonal cleared.
__global__ __launch_bounds__(SC_BT) void hessenberg_kernel(double *__restrict__ h, unsigned n, double *__restrict__ tau,
		double *__restrict__ z)
{
	__shared__ double red[SC_BT / 64];
	__shared__ double bc[2];
	const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = SC_BT / 64;
	for(unsigned j = 0; j + 2 < n; j++){
		double *hj = h + (size_t)j * n;       // column j
		double s = 0;
		for(unsigned r = j + 2 + tid; r < n; r += SC_BT) s = fma(hj[r], hj[r], s);
#pragma unroll
		for(int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
		if(lane == 0) red[wave] = s;
		__syncthreads();
		if(tid == 0){
			double sig = 0;
			for(unsigned w = 0; w < nw; w++) sig += red[w];
			const double alpha = hj[j + 1];
			double t = 0, scale = 0, beta = alpha;
			if(sig > 0){
				const double nrm = sqrt(fma(alpha, alpha, sig));
				beta = alpha >= 0 ? -nrm : nrm;
				t = (beta - alpha) / beta;
				scale = 1.0 / (alpha - beta);
			}
			tau[j] = t;
			hj[j + 1] = beta;
			bc[0] = t;
			bc[1] = scale;
		}
		__syncthreads();
		const double t = bc[0], scale = bc[1];
		if(t == 0) continue;                  // every thread sees the same t
		for(unsigned r = j + 2 + tid; r < n; r += SC_BT) hj[r] *= scale;
		__syncthreads();
		// left: columns c > j, rows j+1.. : w = v^T h_c, h_c -= t w v   (one wave per column)
		for(unsigned c = j + 1 + wave; c < n; c += nw){
			double *hc = h + (size_t)c * n;
			double w = 0;
			for(unsigned r = j + 2 + lane; r < n; r += 64) w = fma(hj[r], hc[r], w);
#pragma unroll
			for(int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
			w += hc[j + 1];
			const double tw = t * w;
			if(lane == 0) hc[j + 1] -= tw;
			for(unsigned r = j + 2 + lane; r < n; r += 64) hc[r] = fma(-tw, hj[r], hc[r]);
		}
		__syncthreads();
		// right: every row r, columns j+1.. : w = h_r. v, h_r. -= t w v^T   (one thread per row)
		for(unsigned r = tid; r < n; r += SC_BT){
			double w = h[(size_t)(j + 1) * n + r];
			for(unsigned c = j + 2; c < n; c++) w = fma(h[(size_t)c * n + r], hj[c], w);
			const double tw = t * w;
			h[(size_t)(j + 1) * n + r] -= tw;
			for(unsigned c = j + 2; c < n; c++) h[(size_t)c * n + r] = fma(-tw, hj[c], h[(size_t)c * n + r]);
		}
		__syncthreads();
	}
	// Z = H_0 ... H_{n-3} I (backward; reflector j acts on rows and columns j+1..)
	for(size_t e = tid; e < (size_t)n * n; e += SC_BT){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		z[e] = r == c ? 1.0 : 0.0;
	}
	__syncthreads();
	for(int j = (int)n - 3; j >= 0; j--){
		const double t = tau[j];
		if(t != 0){
			const double *v = h + (size_t)j * n;
			for(unsigned c = (unsigned)j + 1 + wave; c < n; c += nw){
				double *zc = z + (size_t)c * n;
				double w = 0;
				for(unsigned r = j + 2 + lane; r < n; r += 64) w = fma(v[r], zc[r], w);
#pragma unroll
				for(int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
				w += zc[j + 1];
				const double tw = t * w;
				if(lane == 0) zc[j + 1] -= tw;
				for(unsigned r = j + 2 + lane; r < n; r += 64) zc[r] = fma(-tw, v[r], zc[r]);
			}
		}
		__syncthreads();
	}
	for(size_t e = tid; e < (size_t)n * n; e += SC_BT){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		if(r > c + 1) h[e] = 0;
	}
}

// Standardisation of the 2 x 2 block [a b; c d] by a rotation R = [cs -sn; sn cs]:
// [a b; c d] = R [a' b'; c' d'] R^T with the block upper triangular for a real pair and
// a' = d', b' c' < 0 for a complex pair (the form of LAPACK's dlanv2, derived from the
// eigenvalues of the block).  Eigenvalues returned in (rt1r, rt1i), (rt2r, rt2i).
__device__ void standardize2(double &a, double &b, double &c, double &d, double &cs, double &sn,
		double &rt1r, double &rt1i, double &rt2r, double &rt2i)
{
	const double eps = 2.220446049250313e-16;
	cs = 1; sn = 0;
	if(c == 0){
	}else if(b == 0){
		// swap rows and columns
		cs = 0; sn = 1;
		const double tmp = d;
		d = a; a = tmp;
		b = -c; c = 0;
	}else if(a - d == 0 && (b >= 0) != (c >= 0)){
		// already standard (complex pair)
	}else{
		const double temp = a - d;
		double p = 0.5 * temp;
		const double bcmax = fmax(fabs(b), fabs(c));
		const double bcmis = fmin(fabs(b), fabs(c)) * hsign(1.0, b) * hsign(1.0, c);
		const double scale = fmax(fabs(p), bcmax);
		double zz = p / scale * p + bcmax / scale * bcmis;
		if(zz >= 4 * eps){
			// real eigenvalues: make the block upper triangular
			zz = p + hsign(sqrt(scale) * sqrt(zz), p);
			a = d + zz;
			d = d - bcmax / zz * bcmis;
			const double tau = hypot(c, zz);
			cs = zz / tau;
			sn = c / tau;
			b = b - c;
			c = 0;
		}else{
			// complex (or nearly equal real) eigenvalues: equal diagonal entries
			const double sigma = b + c;
			const double tau = hypot(sigma, temp);
			cs = sqrt(0.5 * (1 + fabs(sigma) / tau));
			sn = -(p / (tau * cs)) * hsign(1.0, sigma);
			const double aa = a * cs + b * sn, bb = -a * sn + b * cs;
			const double cc = c * cs + d * sn, dd = -c * sn + d * cs;
			a = aa * cs + cc * sn;
			b = bb * cs + dd * sn;
			c = -aa * sn + cc * cs;
			d = -bb * sn + dd * cs;
			const double mid = 0.5 * (a + d);
			a = mid;
			d = mid;
			if(c != 0){
				if(b != 0){
					if((b >= 0) == (c >= 0)){
						// real eigenvalues after all: reduce to upper triangular
						const double sab = sqrt(fabs(b)), sac = sqrt(fabs(c));
						p = hsign(sab * sac, c);
						const double tt = 1 / sqrt(fabs(b + c));
						a = mid + p;
						d = mid - p;
						b = b - c;
						c = 0;
						const double cs1 = sab * tt, sn1 = sac * tt;
						const double t2 = cs * cs1 - sn * sn1;
						sn = cs * sn1 + sn * cs1;
						cs = t2;
					}
				}else{
					b = -c;
					c = 0;
					const double t2 = cs;
					cs = -sn;
					sn = t2;
				}
			}
		}
	}
	rt1r = a;
	rt2r = d;
	if(c == 0){
		rt1i = 0;
		rt2i = 0;
	}else{
		rt1i = sqrt(fabs(b)) * sqrt(fabs(c));
		rt2i = -rt1i;
	}
}

// Francis double-shift QR on the Hessenberg h (column-major) with the transformations accumulated
// into z; h becomes quasi-upper-triangular (2 x 2 blocks standardised), wr/wi the eigenvalues.
// status: 0, or 1 + the index that did not converge within 30 n sweeps.
struct FrancisCtl {
	int nn, l, its, total, k, phase;
	double x, y, zz, q, r;        // reflector of the current chase step
	int active;                   // the step applies a reflector
	int three;                    // 3-element reflector (else 2)
};

__device__ __forceinline__ double &H(double *h, unsigned n, int r, int c){ return h[(size_t)c * n + r]; }

__global__ __launch_bounds__(SC_BT) void francis_kernel(double *__restrict__ h, double *__restrict__ z, unsigned n,
		double *__restrict__ wr, double *__restrict__ wi, int *__restrict__ status)
{
	__shared__ FrancisCtl ctl;
	__shared__ int lmax;
	__shared__ double rot[2];
	const int tid = threadIdx.x;
	const double eps = 2.220446049250313e-16;
	// a scale for all-zero neighbourhoods: the largest |entry| of the Hessenberg matrix
	__shared__ double hmax;
	__shared__ int nonfinite;
	if(tid == 0){ ctl.nn = (int)n - 1; ctl.its = 0; ctl.total = 0; hmax = 0; nonfinite = 0; }
	__syncthreads();
	{
		double m = 0;
		int bad = 0;
		for(size_t e = tid; e < (size_t)n * n; e += SC_BT){
			const double v = fabs(h[e]);
			bad |= !(v <= DBL_MAX);           // NaN or infinite
			m = fmax(m, v);
		}
		// wave maximum, then the first lane of each wave folds into hmax
#pragma unroll
		for(int off = 32; off > 0; off >>= 1){
			m = fmax(m, __shfl_xor(m, off, 64));
			bad |= __shfl_xor(bad, off, 64);
		}
		if((tid & 63) == 0){
			atomicMax((unsigned long long *)&hmax, __double_as_longlong(m));   // m >= 0: order-preserving bits
			if(bad) atomicOr(&nonfinite, 1);
		}
	}
	__syncthreads();
	// a non-finite entry never deflates: stop at once (status n, NaN eigenvalues) instead of
	// running the 30 n sweeps of the cap
	if(nonfinite){
		for(unsigned i = tid; i < n; i += SC_BT){
			wr[i] = __longlong_as_double(0x7ff8000000000000LL);
			wi[i] = 0;
		}
		if(tid == 0) *status = (int)n;
		return;
	}
	const double hscale = hmax > 0 ? hmax : 1.0;
	while(true){
		__syncthreads();
		const int nn = ctl.nn;
		if(nn < 0) break;
		// deflation: the largest k in [1, nn] with a negligible subdiagonal entry (0 if none)
		if(tid == 0) lmax = 0;
		__syncthreads();
		for(int k = tid + 1; k <= nn; k += SC_BT){
			double s = fabs(H(h, n, k - 1, k - 1)) + fabs(H(h, n, k, k));
			if(s == 0) s = hscale;
			if(fabs(H(h, n, k, k - 1)) <= eps * s) atomicMax(&lmax, k);
		}
		__syncthreads();
		const int l = lmax;
		if(l > 0 && tid == 0) H(h, n, l, l - 1) = 0;
		if(l == nn){
			if(tid == 0){
				wr[nn] = H(h, n, nn, nn);
				wi[nn] = 0;
				ctl.nn = nn - 1;
				ctl.its = 0;
			}
			continue;
		}
		if(l == nn - 1){
			// a 2 x 2 block: standardise it and carry the rotation through the rest of h and z
			if(tid == 0){
				double a = H(h, n, nn - 1, nn - 1), b = H(h, n, nn - 1, nn), c = H(h, n, nn, nn - 1), d = H(h, n, nn, nn);
				double cs, sn, r1r, r1i, r2r, r2i;
				standardize2(a, b, c, d, cs, sn, r1r, r1i, r2r, r2i);
				H(h, n, nn - 1, nn - 1) = a; H(h, n, nn - 1, nn) = b; H(h, n, nn, nn - 1) = c; H(h, n, nn, nn) = d;
				wr[nn - 1] = r1r; wi[nn - 1] = r1i; wr[nn] = r2r; wi[nn] = r2i;
				rot[0] = cs;
				rot[1] = sn;
			}
			__syncthreads();
			const double cs = rot[0], sn = rot[1];
			// rows nn-1, nn over the columns right of the block: R^T [x; y]
			for(int j = nn + 1 + tid; j < (int)n; j += SC_BT){
				const double x = H(h, n, nn - 1, j), y = H(h, n, nn, j);
				H(h, n, nn - 1, j) = cs * x + sn * y;
				H(h, n, nn, j) = -sn * x + cs * y;
			}
			// columns nn-1, nn over the rows above the block, and of z: [x y] R
			for(int i = tid; i < nn - 1; i += SC_BT){
				const double x = H(h, n, i, nn - 1), y = H(h, n, i, nn);
				H(h, n, i, nn - 1) = cs * x + sn * y;
				H(h, n, i, nn) = -sn * x + cs * y;
			}
			for(int i = tid; i < (int)n; i += SC_BT){
				const double x = z[(size_t)(nn - 1) * n + i], y = z[(size_t)nn * n + i];
				z[(size_t)(nn - 1) * n + i] = cs * x + sn * y;
				z[(size_t)nn * n + i] = -sn * x + cs * y;
			}
			__syncthreads();
			if(tid == 0){ ctl.nn = nn - 2; ctl.its = 0; }
			continue;
		}
		// a double-shift sweep over [l, nn]
		if(tid == 0){
			ctl.its++;
			ctl.total++;
			if(ctl.total > 30 * (int)n){ *status = nn + 1; ctl.nn = -1; }
		}
		__syncthreads();
		if(ctl.nn < 0) break;
		for(int k = l; k <= nn - 1; k++){
			if(tid == 0){
				double p, q, r;
				if(k == l){
					// first column of (H - s1)(H - s2), shifts from the trailing 2 x 2 block (an ad hoc
					// shift at sweeps 10 and 20 of one eigenvalue)
					double x = H(h, n, nn, nn), y = H(h, n, nn - 1, nn - 1), w = H(h, n, nn, nn - 1) * H(h, n, nn - 1, nn);
					if(ctl.its == 10 || ctl.its == 20){
						const double s = fabs(H(h, n, nn, nn - 1)) + fabs(H(h, n, nn - 1, nn - 2));
						x = y = 0.75 * s + H(h, n, nn, nn);
						w = -0.4375 * s * s;
					}
					const double zl = H(h, n, l, l), rr = x - zl, ss = y - zl;
					p = (rr * ss - w) / H(h, n, l + 1, l) + H(h, n, l, l + 1);
					q = H(h, n, l + 1, l + 1) - zl - rr - ss;
					r = H(h, n, l + 2, l + 1);
					const double sc = fabs(p) + fabs(q) + fabs(r);
					p /= sc; q /= sc; r /= sc;
					ctl.x = 1;      // scale of the first column (its own entry h(l, l-1) is not touched)
				}else{
					p = H(h, n, k, k - 1);
					q = H(h, n, k + 1, k - 1);
					r = k != nn - 1 ? H(h, n, k + 2, k - 1) : 0.0;
					const double sc = fabs(p) + fabs(q) + fabs(r);
					ctl.x = sc;
					if(sc != 0){ p /= sc; q /= sc; r /= sc; }
				}
				const double s = hsign(sqrt(p * p + q * q + r * r), p);
				ctl.active = 0;
				if(ctl.x != 0 && s != 0){
					if(k != l){
						H(h, n, k, k - 1) = -s * ctl.x;
						H(h, n, k + 1, k - 1) = 0;
						if(k != nn - 1) H(h, n, k + 2, k - 1) = 0;
					}
					p += s;
					ctl.active = 1;
					ctl.three = k != nn - 1;
					ctl.x = p / s;
					ctl.y = q / s;
					ctl.zz = r / s;
					ctl.q = q / p;
					ctl.r = r / p;
				}
			}
			__syncthreads();
			if(ctl.active){
				const double x = ctl.x, y = ctl.y, zz = ctl.zz, q = ctl.q, r = ctl.r;
				const bool three = ctl.three != 0;
				// rows k, k+1 (, k+2), columns k .. n-1
				for(int j = k + tid; j < (int)n; j += SC_BT){
					double p = H(h, n, k, j) + q * H(h, n, k + 1, j);
					if(three){
						p += r * H(h, n, k + 2, j);
						H(h, n, k + 2, j) -= p * zz;
					}
					H(h, n, k, j) -= p * x;
					H(h, n, k + 1, j) -= p * y;
				}
				__syncthreads();
				// columns k, k+1 (, k+2), rows 0 .. min(nn, k+3); the same columns of z
				const int rmax = min(nn, k + 3);
				for(int i = tid; i <= rmax; i += SC_BT){
					double p = x * H(h, n, i, k) + y * H(h, n, i, k + 1);
					if(three){
						p += zz * H(h, n, i, k + 2);
						H(h, n, i, k + 2) -= p * r;
					}
					H(h, n, i, k) -= p;
					H(h, n, i, k + 1) -= p * q;
				}
				for(int i = tid; i < (int)n; i += SC_BT){
					double *z0 = z + (size_t)k * n, *z1 = z0 + n;
					double p = x * z0[i] + y * z1[i];
					if(three){
						p += zz * z1[n + i];
						z1[n + i] -= p * r;
					}
					z0[i] -= p;
					z1[i] -= p * q;
				}
			}
			__syncthreads();
		}
	}
	__syncthreads();
	// quasi-triangular: clear what lies below the subdiagonal
	for(size_t e = tid; e < (size_t)n * n; e += SC_BT){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		if(r > c + 1) h[e] = 0;
	}
}


// B = M_hat - 1 into h (column-major copy)
__global__ void minus_identity_kernel(const double *__restrict__ m, double *__restrict__ h, unsigned n)
{
	const size_t total = (size_t)n * n;
	for(size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		h[e] = m[e] - (r == c ? 1.0 : 0.0);
	}
}

// organic_dense_matrix.c:562-576: min / max real part of the eigenvalues, the threshold
// eps = max_ev DBL_EPSILON / 1e-3, and when min_ev < -1 or eps > 1 every entry of T with |T_ij| < eps
// set to zero; then 1 + T.  One workgroup.
__global__ __launch_bounds__(1024) void schur_threshold_kernel(double *__restrict__ t, const double *__restrict__ wr,
		unsigned n, double *__restrict__ info)
{
	__shared__ double mn[16], mx[16];
	const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	double lo = wr[0], hi = wr[0];
	for(unsigned i = tid; i < n; i += 1024){ lo = fmin(lo, wr[i]); hi = fmax(hi, wr[i]); }
#pragma unroll
	for(int off = 32; off > 0; off >>= 1){
		lo = fmin(lo, __shfl_xor(lo, off, 64));
		hi = fmax(hi, __shfl_xor(hi, off, 64));
	}
	if(lane == 0){ mn[wave] = lo; mx[wave] = hi; }
	__syncthreads();
	lo = mn[0]; hi = mx[0];
	for(int w = 1; w < 16; w++){ lo = fmin(lo, mn[w]); hi = fmax(hi, mx[w]); }
	const double eps = hi * 2.220446049250313e-16 / 1e-3;
	const bool zero = lo < -1 || eps > 1;
	for(size_t e = tid; e < (size_t)n * n; e += 1024){
		const unsigned c = (unsigned)(e / n), r = (unsigned)(e - (size_t)c * n);
		double v = t[e];
		if(zero && fabs(v) < eps) v = 0;
		t[e] = v + (r == c ? 1.0 : 0.0);
	}
	if(tid == 0){
		info[0] = lo;
		info[1] = hi;
		info[2] = zero ? eps : 0.0;
	}
}

// out = A op(B), column-major N x N, op(B) = B^T (TB) or B: 64 x 64 tiles, 256 threads of 4 x 4
// outputs, k in panels of 16 through LDS
template <bool TB>
__global__ __launch_bounds__(256) void schur_gemm_kernel(const double *__restrict__ a, const double *__restrict__ b,
		unsigned n, double *__restrict__ out)
{
	__shared__ double as[16][65], bs[16][65];
	const unsigned r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
	const unsigned tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
	double acc[4][4] = {};
	for(unsigned k0 = 0; k0 < n; k0 += 16){
		for(unsigned e = threadIdx.x; e < 16 * 64; e += 256){
			{
				const unsigned kk = e >> 6, i = e & 63, k = k0 + kk;
				as[kk][i] = (k < n && r0 + i < n) ? a[(size_t)k * n + r0 + i] : 0.0;
			}
			if(TB){
				const unsigned kk = e >> 6, j = e & 63, k = k0 + kk;      // B^T(k, j) = B(j, k)
				bs[kk][j] = (k < n && c0 + j < n) ? b[(size_t)k * n + c0 + j] : 0.0;
			}else{
				const unsigned kk = e & 15, j = e >> 4, k = k0 + kk;      // B(k, j): contiguous in k
				bs[kk][j] = (k < n && c0 + j < n) ? b[(size_t)(c0 + j) * n + k] : 0.0;
			}
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

// a failed Schur iteration leaves log det NaN (a trajectory built on it is rejected)
__global__ void schur_status_kernel(const int *__restrict__ status, double *__restrict__ scal)
{
	if(*status != 0){
		scal[SCAL_LOGDET] = __longlong_as_double(0x7ff8000000000000LL);
		scal[SCAL_SCHUR_FAIL] = *status;
	}
}

int schur_reserve(ephmc_ctx *c)
{
	if(c->schur.H) return EPHMC_OK;
	const size_t mat = c->mat, n = c->ns;
	double **d[] = {&c->schur.H, &c->schur.Z, &c->schur.P, &c->schur.tau, &c->schur.wr, &c->schur.wi, &c->schur.info};
	const size_t sz[] = {mat, mat, mat, n, n, n, 4};
	for(int i = 0; i < 7; i++){
		if(hipMalloc((void **)d[i], sz[i] * sizeof(double)) != hipSuccess){
			(void)hipGetLastError();
			schur_release(c);
			set_error("Schur form: device buffers (%zu MB) could not be allocated", (3 * mat * 8) >> 20);
			return EPHMC_ERR_ALLOC;
		}
	}
	if(hipMalloc((void **)&c->schur.status, sizeof(int)) != hipSuccess){
		(void)hipGetLastError();
		schur_release(c);
		set_error("Schur form: device buffers could not be allocated");
		return EPHMC_ERR_ALLOC;
	}
	return EPHMC_OK;
}

void schur_release(ephmc_ctx *c)
{
	void *b[] = {c->schur.H, c->schur.Z, c->schur.P, c->schur.tau, c->schur.wr, c->schur.wi, c->schur.info, c->schur.status};
	for(void *p : b) if(p) (void)hipFree(p);
	c->schur = SchurWork{};
}

__global__ void zero_int_kernel(int *p){ *p = 0; }

// M <- M^{-1} through the real Schur form of M - 1 (SCHUR_FORM_INV), log det into scal.
int schur_logdet_inverse(ephmc_ctx *c, double *M)
{
	const unsigned n = c->ns;
	if(n > 2048){ set_error("Schur form: N = %u exceeds 2048", n); return EPHMC_ERR_UNSUPPORTED; }
	EPHMC_TRY(schur_reserve(c));
	SchurWork &w = c->schur;
	const size_t total = (size_t)n * n;
	minus_identity_kernel<<<(unsigned)std::min<size_t>((total + 255) / 256, 4096), 256, 0, c->stream>>>(M, w.H, n);
	EPHMC_LAUNCHED();
	zero_int_kernel<<<1, 1, 0, c->stream>>>(w.status);
	EPHMC_LAUNCHED();
	hessenberg_kernel<<<1, SC_BT, 0, c->stream>>>(w.H, n, w.tau, w.Z);
	EPHMC_LAUNCHED();
	francis_kernel<<<1, SC_BT, 0, c->stream>>>(w.H, w.Z, n, w.wr, w.wi, w.status);
	EPHMC_LAUNCHED();
	schur_threshold_kernel<<<1, 1024, 0, c->stream>>>(w.H, w.wr, n, w.info);
	EPHMC_LAUNCHED();
	EPHMC_TRY(gj_logdet_inverse(c, w.H));                 // (1 + T)^{-1}, log|det(1 + T)|
	const dim3 gg((n + 63) / 64, (n + 63) / 64);
	schur_gemm_kernel<true><<<gg, 256, 0, c->stream>>>(w.H, w.Z, n, w.P);     // (1 + T)^{-1} Z^T
	EPHMC_LAUNCHED();
	schur_gemm_kernel<false><<<gg, 256, 0, c->stream>>>(w.Z, w.P, n, M);      // Z (1 + T)^{-1} Z^T
	EPHMC_LAUNCHED();
	schur_status_kernel<<<1, 1, 0, c->stream>>>(w.status, c->d_scal);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// CONSTANT_FORCE under SCHUR_FORM_INV (organic_dense_matrix.c:270-278 and organic_hmc.c:100-126):
// QM1P = 1 - M_hat^{-1} for every slice, projection Q[i,nb] + Q[nb,i] clipped to [-2, 2]
__global__ void schur_constant_force_kernel(const double *__restrict__ minv, const int *__restrict__ nb, unsigned ns,
		double f0, double f1, double f2, double *__restrict__ pdot)
{
	const double f[3] = {f0, f1, f2};
	for(unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < 3 * ns; e += gridDim.x * blockDim.x){
		const unsigned i = e / 3, k = e - 3 * i;
		const unsigned j = (unsigned)nb[(size_t)(2 * k) * ns + i];
		double proj = -(minv[(size_t)j * ns + i] + minv[(size_t)i * ns + j]) + (i == j ? 2.0 : 0.0);
		if(fabs(proj) > 2) proj = proj > 0 ? 2.0 : -2.0;
		pdot[e] = f[k] * proj;
	}
}

int schur_constant_force(ephmc_ctx *c, const double *minv, double *pdot)
{
	const double f[3] = {-c->cfg.kappa[0] * c->cfg.flavors, -c->cfg.kappa[1] * c->cfg.flavors, -c->cfg.kappa[2] * c->cfg.flavors};
	schur_constant_force_kernel<<<(3 * c->ns + 255) / 256, 256, 0, c->stream>>>(minv, c->d_nb, c->ns, f[0], f[1], f[2], pdot);
	EPHMC_LAUNCHED();
	return broadcast_slice0(c, pdot);
}

} // namespace ephmc

using namespace ephmc;

extern "C" int ephmc_schur(ephmc_ctx *c, const double *B, double *Z, double *T, double *wr, double *wi)
{
	if(!c || !B || !Z || !T){ set_error("null argument"); return EPHMC_ERR_ARG; }
	const unsigned n = c->ns;
	if(n > 2048){ set_error("Schur form: N = %u exceeds 2048", n); return EPHMC_ERR_UNSUPPORTED; }
	EPHMC_TRY(schur_reserve(c));
	SchurWork &w = c->schur;
	EPHMC_HIP(hipMemcpyAsync(T, B, c->mat * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	zero_int_kernel<<<1, 1, 0, c->stream>>>(w.status);
	EPHMC_LAUNCHED();
	hessenberg_kernel<<<1, SC_BT, 0, c->stream>>>(T, n, w.tau, Z);
	EPHMC_LAUNCHED();
	francis_kernel<<<1, SC_BT, 0, c->stream>>>(T, Z, n, w.wr, w.wi, w.status);
	EPHMC_LAUNCHED();
	int st = 0;
	EPHMC_HIP(hipMemcpyAsync(&st, w.status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
	if(wr) EPHMC_HIP(hipMemcpyAsync(wr, w.wr, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
	if(wi) EPHMC_HIP(hipMemcpyAsync(wi, w.wi, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	if(st != 0){ set_error("Schur form: the QR iteration did not converge at index %d", st - 1); return EPHMC_ERR_NUMERIC; }
	return EPHMC_OK;
}
